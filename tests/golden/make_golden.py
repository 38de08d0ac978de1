"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in the build container after `make -C oracle ref` (the reference is
compiled from its own sources into oracle/_ref/libohv1_ref_serial.so; it is
never shipped).  Outputs are data only:

* c1_states.npz -- every array of the mesh after each step of the C1
  workload (3D Kuhn box 8^3, h = 0.05: refine, coarsen, swap);
* workload_digests.json -- per step of each scripted workload: the pass
  return value, entity counts and a SHA-256 of every array of the state;
* known_answers.json -- the reference's swap interior-edge tables,
  choose_edge_swap on seeded random rings and quality values of named
  simplices.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import ctypes as C  # noqa: E402

import meshlib as ML  # noqa: E402
import parity_cases as P  # noqa: E402

DIGEST_CASES = ["c1", "warp2d_sine_n16", "warp2d_rot_n32", "warp3d_rot_n4"]


def state_digest(m):
    st = m.state()
    per = {}
    for k in sorted(st):
        v = st[k]
        if isinstance(v, int):
            per[k] = v
        else:
            per[k] = hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
    return per


def ring_cases(seed=7, per_size=40):
    """Seeded random rings around the z axis (edge from z=-1 to z=+1)."""
    rng = np.random.default_rng(seed)
    out = []
    for rs in range(3, 8):
        for _ in range(per_size):
            ang = np.sort(rng.uniform(0, 2 * np.pi, rs))
            rad = rng.uniform(0.3, 1.5, rs)
            ring = np.stack([rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-0.4, 0.4, rs)], 1)
            edge = np.array([[0.0, 0.0, -1.0], [0.0, 0.0, 1.0]]) + rng.normal(0, 0.05, (2, 3))
            out.append((rs, edge, ring))
    return out


def main():
    ref = ML.load("ref")
    if ref is None:
        sys.exit("oracle/_ref/libohv1_ref_serial.so is not built (make -C oracle ref)")
    # C1 full arrays
    arrays = {}
    for label, m in P.c1_steps(ref):
        for k, v in m.state().items():
            arrays["%s|%s" % (label, k)] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, "c1_states.npz"), **arrays)
    # digests
    dig = {}
    for case in DIGEST_CASES:
        steps = []
        for label, m in P.CASES[case](ref):
            steps.append({"label": label, "last": getattr(m, "last", None),
                          "counts": m.counts(), "arrays": state_digest(m)})
        dig[case] = steps
    steps = []
    for label, m in P.c2_steps(ref):
        steps.append({"label": label, "last": m.last, "counts": m.counts(), "arrays": state_digest(m)})
    dig["c2_n256"] = steps
    with open(os.path.join(HERE, "workload_digests.json"), "w") as f:
        json.dump(dig, f, indent=1, sort_keys=True)
    # known answers
    d = ref.dll
    d.ref_swap_nint_edges.restype = C.c_uint
    d.ref_choose_edge_swap.restype = C.c_uint
    d.ref_choose_edge_swap.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    for fn in ("ref_tet_quality", "ref_triangle_quality", "ref_triangle_xy_quality"):
        getattr(d, fn).restype = C.c_double
        getattr(d, fn).argtypes = [C.c_void_p]
    counts = {3: 1, 4: 2, 5: 5, 6: 14, 7: 42}
    tables = {}
    for rs in range(4, 8):
        k = d.ref_swap_nint_edges(rs)
        rows = []
        for code in range(counts[rs]):
            buf = (C.c_uint * (2 * k))()
            d.ref_swap_int_edges(rs, code, buf)
            rows.append(list(buf))
        tables[str(rs)] = rows
    swaps = []
    for rs, edge, ring in ring_cases():
        e = np.ascontiguousarray(edge)
        r = np.ascontiguousarray(ring)
        q = C.c_double()
        code = d.ref_choose_edge_swap(rs, e.ctypes.data, r.ctypes.data, C.byref(q))
        swaps.append({"ring_size": rs, "edge": e.ravel().tolist(), "ring": r.ravel().tolist(),
                      "code": int(code), "quality": q.value})
    simplices = {
        "kuhn_tet": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]],
        "regular_tet": [[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]],
        "sliver_tet": [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 1e-3]],
        "inverted_tet": [[0, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, 1]],
        "right_tri": [[0, 0, 0], [1, 0, 0], [1, 1, 0]],
        "equilateral_tri": [[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]],
        "cw_tri": [[0, 0, 0], [0, 1, 0], [1, 0, 0]],
    }
    quals = {}
    for name, xs in simplices.items():
        x = np.ascontiguousarray(np.array(xs, dtype=np.float64))
        if len(xs) == 4:
            quals[name] = {"x": x.ravel().tolist(), "tet": d.ref_tet_quality(x.ctypes.data)}
        else:
            quals[name] = {"x": x.ravel().tolist(), "tri": d.ref_triangle_quality(x.ctypes.data),
                           "tri_xy": d.ref_triangle_xy_quality(x.ctypes.data)}
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump({"swap_int_edges": tables, "choose_edge_swap": swaps, "qualities": quals}, f,
                  indent=1, sort_keys=True)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
