"""GPU known-answer test from the reference's batch prototype
(code/batch_proto/batch_proto.cu:257-409, recorded run run_dispatch.txt:3-16):
level-wise frequent-itemset search over its planted-pattern bitmaps (N = 2^23
transactions, 512 items, minsup 12,582), every level counted by the fused
batched path with each prefix's tail restricted to its canonical extensions
(items j > last).  F1 = 248 and levels 2..7 find 585, 587, 350, 117, 17, 0
(1,904 in all); all 80 planted patterns are recalled."""
import numpy as np
import pytest

import oracle_lib
import tensorfim_amd as tf

pytestmark = pytest.mark.gpu


def test_batch_prototype_levels_on_gpu():
    B, pats = oracle_lib.proto_bitmaps()
    n_bits, minsup = 1 << 23, int((1 << 23) * 0.15 / 100.0)
    with tf.Context(0) as ctx:
        ctx.load_bitmaps(B, n_bits, 512)
        f1 = ctx.item_supports()
        freq1 = [j for j in range(512) if f1[j] >= minsup]
        assert len(freq1) == 248
        cur = [[j] for j in freq1]
        found = {tuple(c) for c in cur}
        levels = []
        for level in range(2, 9):
            if not cur:
                break
            tails = [[j for j in freq1 if j > p[-1]] for p in cur]
            nxt = []
            for i in range(0, len(cur), 1024):   # waves of <= 1024 prefixes (batch_proto BATCH)
                counts = ctx.count_nodes(cur[i:i + 1024], tails[i:i + 1024])
                for p, t, c in zip(cur[i:i + 1024], tails[i:i + 1024], counts):
                    nxt += [p + [j] for j, v in zip(t, c) if v >= minsup]
            levels.append(len(nxt))
            found |= {tuple(x) for x in nxt}
            cur = nxt
            print("level", level, len(nxt), flush=True)
        assert levels == [585, 587, 350, 117, 17, 0]
        assert len(found) == 1904
        assert sum(tuple(sorted(p)) in found for p in pats) == 80
