"""Restatement of the counter-based streams the learned search draws its
proposals from on the GPU (graphgolf-cuda_amd/csrc/gg_devrng.hpp, k_propose in
gg_learn_search.hip), and of the 2-opt proposal attempt they feed.  Test
infrastructure."""
M64 = (1 << 64) - 1


def mix(z):
    """splitmix64 finaliser (Steele, Lea, Flood 2014)."""
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


class Stream:
    def __init__(self, seed=0, step=0, lane=0, raw_state=None):
        if raw_state is not None:
            self.s = raw_state & M64
        else:
            self.s = (mix(mix(seed ^ 0xA0761D6478BD642F) ^ ((step * 0xE7037ED1A0B428DB) & M64))
                      ^ ((lane * 0x8EBC6AF09C88C6E3) & M64))

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        return mix(self.s)

    def below(self, r):
        """Uniform integer in [0, r): Lemire's multiply-shift with rejection."""
        x = self.next()
        m = x * r
        if (m & M64) < r:
            t = ((1 << 64) - r) % r
            while (m & M64) < t:
                x = self.next()
                m = x * r
        return m >> 64


def propose(adj, seed, step, p, max_att=4096):
    """Move row (a, b, c, d, kind) of proposal p of step `step`, or None."""
    n, d = adj.shape
    st = Stream(seed, step, p)
    for _ in range(max_att):
        a, ia = st.below(n), st.below(d)
        c, ic = st.below(n), st.below(d)
        kind = st.below(2)
        b, dd = int(adj[a, ia]), int(adj[c, ic])
        if len({a, b, c, dd}) < 4:
            continue
        xa, xb = (dd, c) if kind else (c, dd)
        if xa in adj[a] or xb in adj[b]:
            continue
        if a not in adj[b] or c not in adj[dd]:
            continue
        return (a, b, c, dd, kind)
    return None
