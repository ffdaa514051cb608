"""The host kd-tree's node layout: the root at node 0, an unreferenced empty
leaf at node 1, every inner node's child pair on an even node (one aligned
16-byte record, which the device traversal loads whole), every node
referenced once, leaf ranges tiling the reference list. Checked for a
one-thread build and for the multi-threaded build, whose subtrees are built
apart and stitched together. Host-only."""
import numpy as np
import pytest

import xtracer_amd as xa
from tests.helpers import scene_path


def _check(nodes, refs):
    word, cnt = nodes[:, 0], nodes[:, 1]
    leaf = (word & 3) == 3
    assert leaf[1] and cnt[1] == 0
    inner = np.flatnonzero(~leaf)
    child = word[inner] >> 2
    assert (child % 2 == 0).all() and child.min() >= 2 and child.max() + 1 < len(nodes)
    seen = np.zeros(len(nodes), np.int64)
    np.add.at(seen, child, 1)
    np.add.at(seen, child + 1, 1)
    assert seen[0] == 0 and seen[1] == 0 and (seen[2:] == 1).all()
    first, n = word[leaf] >> 2, cnt[leaf]
    order = np.lexsort((n, first))  # empty leaves share their offset with the next leaf
    ends = np.cumsum(n[order])
    assert (first[order] == np.concatenate([[0], ends[:-1]])).all() and ends[-1] == len(refs)


@pytest.mark.parametrize("threads", [1, 0])
@pytest.mark.parametrize("kind,kw", [("icosphere", {"subdiv": 6}), ("spheres", {})])
def test_kd_pairs_aligned(kind, kw, threads):
    s = xa.Scene(scene_path(kind, **kw), threads=threads)
    nodes, refs, _ = s.kdtree()
    _check(nodes, refs)
