"""Certified ties between two closest-hit answers.

Pine's BVH::Intersect keeps the first prim its traversal order finds at the
smallest t (bvh.cpp:534-610); this library (and the oracle's brute force)
keeps the lexicographic minimum of (t, prim id), which does not depend on the
tree.  The two may name different prims only where both prims are hit at
exactly the same t.  certify_ties() checks that for every ray whose ids
differ, with the oracle's single-prim intersection tests."""
import numpy as np


def single_prim(tri=None, ana=None):
    """The one-prim argument pair (triangles [k, 9], analytic [m, 10]) of
    oracle.trace_brute for a triangle row [9] or an analytic row [10]."""
    t = np.zeros((0, 9), np.float32) if tri is None else np.asarray(tri, np.float32).reshape(1, 9)
    a = None if ana is None else np.asarray(ana, np.float32).reshape(1, 10)
    return t, a


def mesh_prim_geom(oracle, v, i):
    """prim id -> single_prim() for a single triangle mesh (id = triangle)."""
    tri = oracle.tri9(v, i)
    return lambda pid: single_prim(tri=tri[pid])


def certify_ties(oracle, rays, smaller, other, t_hit, prim_geom):
    """For every ray with smaller[k] != other[k] (both hits): the oracle finds
    each of the two prims alone hit at exactly t_hit[k], and smaller[k] is the
    smaller id.  Returns the number of ties certified."""
    diff = np.flatnonzero((other >= 0) & (smaller != other))
    for k in diff:
        r = np.ascontiguousarray(rays[k: k + 1], np.float32)
        for pid in (int(other[k]), int(smaller[k])):
            tri, ana = prim_geom(pid)
            bp, bt, _, _, _ = oracle.trace_brute(tri, r, ana)
            assert bp[0] == 0 and bt[0].view(np.int32) == np.float32(t_hit[k]).view(np.int32), \
                f"ray {k}: prim {pid} is not hit at t = {t_hit[k]!r}: an uncertified id difference"
        assert smaller[k] < other[k], f"ray {k}: tie resolved to the larger id {smaller[k]}"
    return len(diff)
