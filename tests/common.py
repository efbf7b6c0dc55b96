"""Shared helpers for the parity tests (workload setup for oracle and device)."""
import numpy as np

from mfcl import workloads as W


def hull_shapes_for_oracle(h, idx_hulls):
    """Faces (qhull) for the listed hulls; returns oracle ShapeSet arguments
    covering only those hulls, and the map shape -> local hull id."""
    vs, ts, vo, vc, to, tc = [], [], [], [], [], []
    o = q = 0
    for i in idx_hulls:
        pts = h["verts"][h["vert_off"][i]:h["vert_off"][i] + h["vert_cnt"][i]]
        keep, tri = W.hull_faces(pts)
        v = pts[keep]
        vs.append(v)
        ts.append(tri)
        vo.append(o)
        vc.append(len(v))
        to.append(q)
        tc.append(len(tri))
        o += len(v)
        q += len(tri)
    return (np.concatenate(vs), np.array(vo), np.array(vc), np.concatenate(ts), np.array(to), np.array(tc))


def scale_of(types, params, pa, pb, verts=None, vert_off=None, vert_cnt=None):
    r = W.bounding_radius(types, params, vert_cnt, verts, vert_off)
    return r[pa] + r[pb]
