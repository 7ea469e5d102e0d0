"""Seed-search reference on the oracle (crack/src/gpu_biomes/gpu_biomes.cu:242-266 semantics): per
candidate, the number of observations passed in order; found = candidates passing all."""
import numpy as np

import oracle as O


def obs_quarts(seed, obs):
    """Quart point of every observation for a seed (block observations through BiomeManager zoom)."""
    hs = None
    q = []
    for (x, y, z, block, _mask) in obs:
        if block:
            if hs is None:
                hs = O.hashed_seed(seed)
            q.append(O.zoom(hs, x, y, z))
        else:
            q.append((x, y, z))
    return np.array(q, np.int32)


def search(orc, seeds, obs, ties=False):
    """obs: [(x, y, z, block, set_of_biome_ids)].  Returns (found sorted as unsigned, reached)."""
    reached = np.zeros(len(obs) + 1, np.uint64)
    found = []
    for s in seeds:
        s = int(s)
        pts = obs_quarts(s, obs)
        b, tg = orc.points([s], pts, want_tg=True)
        j = 0
        for i, (_x, _y, _z, _blk, mask) in enumerate(obs):
            ok = int(b[0, i]) in mask
            if not ok and ties and orc.dim != "the_end":
                cnt, m = orc.tie_mask(tg[0, i])
                ok = any((m[bb >> 6] >> (bb & 63)) & 1 for bb in mask)
            if not ok:
                break
            j += 1
        reached[j] += 1
        if j == len(obs):
            found.append(s)
    found.sort(key=lambda v: v & ((1 << 64) - 1))
    return np.array(found, np.int64), reached
