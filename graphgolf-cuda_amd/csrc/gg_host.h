// gg_host.h -- host-side helpers exported by the library (no device work).
#pragma once
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int gg_gen_random_regular(int n, int d, uint64_t seed, int32_t *adj_out);
int gg_gen_orbit_part(int g, int m, int d, uint64_t seed, int32_t *table_out);
#ifdef __cplusplus
}
#endif
