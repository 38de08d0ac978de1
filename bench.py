"""Benchmark of the MI355X adapt hot path.

Default workload (c3): FP32-mixed vs FP64 elements/sec per adapt pass at
equal min quality, on one GPU.
  Mesh: 3D Kuhn box 64^3 cubes x 6 = 1,572,864 tets, generated on the GPU.
  Field: per cycle, the metric from the analytic Hessian of
  f = tanh(50 (x + y + z - 1.5)) (evaluated on the host in FP64, as the
  reference evaluates analytic fields) through the FP64 eigendecomposition,
  eps 0.01, sizes clamped to [0.002, 0.1].
  Run: 8 cycles of mesh_adapt (size floor 1/3, good quality 0.3, 4 sliver
  layers).  Every cycle runs to completion: the operation limit is set to
  1000, above the most any cycle needs (the eight cycles take 95 to 872
  operations; the counts are reported and a cycle at the limit is an
  error), so no cycle stops where the reference's mesh_adapt would abort
  (adapt.cpp:36-45).  The independent sets of this workload need more
  rounds than the reference's 100 (indset.cpp:33): its ties in goodness
  form long chains; the round limit is raised to 2000 (an extension; the
  rounds the workload needed are reported).

  A "step" is one mesh_adapt cycle.  One untimed survey run per precision
  adapts the whole 8-cycle sequence from a fresh box with every kernel timed
  (per-kernel table, dominant kernel) and keeps each cycle's input mesh on
  the GPU: FP32-mixed and FP64 must give the same result, entity counts and
  min quality after every cycle and the identical final mesh (SHA-256 of
  every array).  Then, per precision, W untimed and K
  timed steps adapt copies of the cycle inputs in order (cycle k mod 8; the
  copies are made before the clock starts, so inputs are resident in HBM),
  and each step's return value, pass count and entity counts are checked
  against the survey's.  value = sum over the timed steps of the input
  element counts of the passes that changed the mesh / wall time of the
  timed region (FP32-mixed); fp64_value likewise for the FP64 path.

  roofline: the dominant kernel of the FP32-mixed survey; in the mixed timed
  region only that kernel carries start/stop events (in its own dispatch),
  which give its average duration.
  cpu_baseline: the reference's OpenMP loop build (compiled from its sources
  by oracle/Makefile) on the host's cores, on the isotropic stand-in of this
  workload (v1 cannot express a metric: adapt_size from v1's
  size_from_hessian of the same Hessian, min 0.002, max 0.1), a bounded
  sample of its first passes; the product's FP64 path runs the same
  stand-in on the GPU for comparison and the two final meshes are compared.

Other workloads (--workload), each printing one JSON line of its own:
  c2  2D Kuhn square 256 x 256 x 2 = 131,072 triangles, warped by
      u = 0.1 sin(2 pi x) cos(2 pi y) through warp_to_limit(0.2), then one
      mesh_adapt with h = 1/200 (the FP64 drop-in); a step is one adapt.
  c4  256 independent meshes (numpy PCG64 seed 2024), n log-uniform in
      [8, 32], random rotation, metric axis lengths log-uniform in
      [0.005, 0.1]; one mesh_adapt each, one mesh at a time and
      concurrently (ohx_adapt_many).
  c5  3D Kuhn box 160^3 x 6 = 24,576,000 tets, rotating anisotropic metric
      (short axis 1e-3 in the xy plane at angle 2 pi z, 20:1): the first
      passes of the first cycle, then one swap pass, for per-kernel GB/s at
      the real shape and the peak device memory.
  --calibrate-copy runs plain device-to-device copies of known size, for
  scaling the FETCH_SIZE / WRITE_SIZE counters of a profiler run.
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from omega_h_v1_amd import seam as S  # noqa: E402
from omega_h_v1_amd import workloads as W  # noqa: E402

_LIBC = C.CDLL(None)
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s

# c3
C3_N, C3_CYCLES = 64, 8
C3_A, C3_S0, C3_EPS, C3_HMIN, C3_HMAX = 50.0, 1.5, 0.01, 0.002, 0.1
C3_MAX_OPS = 1000
C3_INDSET_ROUNDS = 2000
ADAPT3 = (1.0 / 3.0, 0.3, 4)
# isotropic stand-in of c3 for the CPU baseline: operations of the bounded
# sample (the reference's adapt returns at the limit in the harness)
C3_CPU_OPS = 7

# c2
N_CELLS = 256
H = 1.0 / 200.0
WARP_FLOOR = 0.2
ADAPT_ARGS = (1.0 / 3.0, 0.3, 4, 50)

# Bytes per launch beyond L2 measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (one run each), scaled by the same counters read for a copy
# kernel of known size (FETCH_SIZE x 2, WRITE_SIZE x 1), averaged over the
# profiled run's launches of the kernel; per workload and kernel (the
# kernels that can be a workload's dominant one)
TRAFFIC = {
    ("c2", "scan"): 2764823,
    ("c2", "indset_round"): 3225804,
    ("c2", "reflect_down"): 53090816,
    ("c3", "coarsen_quality_mixed"): 610625883,
    ("c3", "scan"): 12556348,
    ("c3", "reflect_down"): 715727174,
    ("c5", "coarsen_outputs"): 4469509427,
    ("c5", "up_patch_fill"): 3443524823,
    ("c5", "scan"): 389051928,
    ("c5", "coarsen_quality_mixed"): 13300132659,
    ("c5", "lows_hash_insert"): 2519497975,
    ("c5", "reflect_down"): 11939056067,
}


def _checker(which):
    """The reference's own builds (compiled from its sources into oracle/_ref
    by oracle/Makefile): the CPU baseline's libraries, never the product."""
    path = os.path.join(ROOT, "oracle", "_ref", {"ref": "libohv1_ref_serial.so",
                                                 "ref_omp": "libohv1_ref_omp.so"}[which])
    return S.Lib(path, "ref_") if os.path.exists(path) else None


def digest(m):
    h = hashlib.sha256()
    for k, v in sorted(m.state().items()):
        h.update(k.encode())
        h.update(np.asarray(v).tobytes())
    return h.hexdigest()


def light_digest(m):
    """counts, every dimension's vertex connectivity and the coordinates"""
    h = hashlib.sha256()
    h.update(np.asarray(m.counts(), dtype=np.int64).tobytes())
    for d in range(1, m.dim + 1):
        h.update(m.down(d, 0).tobytes())
    h.update(m.coords().tobytes())
    return h.hexdigest()


# ------------------------------------------------------------ profiling API
def profile_get(dll, name):
    f = dll.ohx_profile_get
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_ulonglong), C.POINTER(C.c_double)]
    ms, n, units = C.c_double(), C.c_ulonglong(), C.c_double()
    f(name.encode(), C.byref(ms), C.byref(n), C.byref(units))
    return ms.value, n.value, units.value


def profile_names(dll):
    f = dll.ohx_profile_names
    f.restype = C.c_uint
    f.argtypes = [C.c_char_p, C.c_uint]
    n = f(None, 0)
    buf = C.create_string_buffer(n + 1)
    f(buf, n + 1)
    return [s for s in buf.value.decode().split("\n") if s]


def tuning_report(dll):
    f = dll.ohx_tuning_report
    f.restype = C.c_uint
    f.argtypes = [C.c_char_p, C.c_uint]
    n = f(None, 0)
    buf = C.create_string_buffer(n + 1)
    f(buf, n + 1)
    return [x for x in buf.value.decode().split("\n") if x]


def setup_api(dll):
    dll.ohx_profile_only.argtypes = [C.c_char_p]
    dll.ohx_profile_only.restype = None
    dll.ohx_set_precision.argtypes = [C.c_int]
    dll.ohx_set_indset_max_rounds.argtypes = [C.c_uint]
    dll.ohx_indset_rounds_max.argtypes = [C.c_int]
    dll.ohx_indset_rounds_max.restype = C.c_uint
    dll.ohx_mixed_stats.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
    dll.ohx_device_bytes_high_water.restype = C.c_ulonglong


def kernel_table(dll):
    tab = {name: profile_get(dll, name) for name in profile_names(dll)}
    host_sync = tab.pop("host_sync", None)
    host_malloc = tab.pop("host_malloc", None)
    return tab, host_sync, host_malloc


def per_kernel(tab):
    return {name: {"ms": round(ms, 3), "launches": n, "avg_us": round(1e3 * ms / max(n, 1), 2),
                   "GB/s": round(nb / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
            for name, (ms, n, nb) in sorted(tab.items(), key=lambda kv: -kv[1][0])}


def roofline_of(dll, dom, workload):
    if not dom:
        return None
    ms, n, nbytes = profile_get(dll, dom)
    if not n or ms <= 0:
        return None
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": TRAFFIC.get((workload, dom)),
            "avg_launch_us": round(1e3 * ms / max(n, 1), 3), "launches": n,
            "algorithmic_bytes_per_launch": round(nbytes / max(n, 1))}


def mixed_stats(dll):
    err, screened, reeval = C.c_double(), C.c_ulonglong(), C.c_ulonglong()
    dll.ohx_mixed_stats(C.byref(err), C.byref(screened), C.byref(reeval))
    return {"max_abs_q32_minus_q64": err.value, "decisions_in_fp32": screened.value,
            "fp64_reevaluations": reeval.value}


# -------------------------------------------------------------------- c3
def c3_set_metric(m):
    x = m.coords().reshape(-1, 3)
    if m.tag(0, "metric", f64=True) is not None:
        m.free_field(0, "metric")
    m.add_field(0, "hessian", W.tanh_front_hessian(x, a=C3_A, s0=C3_S0, dim=3).ravel())
    m.metric_from_hessian("hessian", C3_EPS, C3_HMIN, C3_HMAX)
    m.free_field(0, "hessian")


def c3_survey(lib, mixed):
    """The whole 8-cycle sequence from a fresh box, every kernel timed; keeps
    each cycle's input and hashes each cycle's output."""
    dll = lib.dll
    dll.ohx_set_precision(1 if mixed else 0)
    dll.ohx_mixed_stats_reset()
    dll.ohx_device_reset_high_water()
    m = S.Mesh.box_on_device(lib, 3, C3_N)
    m.derive_model_full()
    dll.ohx_profile_reset()
    dll.ohx_profile_only(None)
    dll.ohx_profile_enable(1)
    cycles, inputs = [], []
    field_s = 0.0
    for _ in range(C3_CYCLES):
        tf = time.perf_counter()
        c3_set_metric(m)
        field_s += time.perf_counter() - tf
        inputs.append(m.clone())
        t0 = time.perf_counter()
        r, st = m.adapt_timed(ADAPT3[0], ADAPT3[1], ADAPT3[2], C3_MAX_OPS)
        wall = time.perf_counter() - t0
        cycles.append({"result": r, "passes": int(st[0]), "elements": float(st[1]), "pass_seconds": st[2],
                       "wall_seconds": wall, "counts": m.counts(), "min_quality": m.min_quality()})
    dll.ohx_synchronize()
    dll.ohx_profile_enable(0)
    at_limit = [k for k, c in enumerate(cycles) if c["result"] < 0]
    if at_limit:
        raise RuntimeError("c3: cycles %r reached the operation limit %d, where the reference's mesh_adapt "
                           "aborts (adapt.cpp:36-45)" % (at_limit, C3_MAX_OPS))
    tab, host_sync, host_malloc = kernel_table(dll)
    dom = max(tab, key=lambda k: tab[k][0]) if tab else None
    out = {
        "cycles": cycles, "inputs": inputs, "kernels": per_kernel(tab), "dominant": dom,
        "host_sync": {"count": host_sync[1], "ms": round(host_sync[0], 3)} if host_sync else None,
        "host_malloc": {"count": host_malloc[1], "ms": round(host_malloc[0], 3)} if host_malloc else None,
        "peak_device_bytes": int(dll.ohx_device_bytes_high_water()),
        "mixed": mixed_stats(dll) if mixed else None,
        "host_field_seconds": field_s,
        "final_min_quality": m.min_quality(),
    }
    # the survey's own rate over the complete 8-cycle run (events on every kernel)
    el = sum(c["elements"] for c in cycles)
    pt = sum(c["pass_seconds"] for c in cycles)
    out["full_run"] = {"elements_over_passes": el, "passes": sum(c["passes"] for c in cycles),
                       "pass_seconds": round(pt, 4), "per_pass_metric": el / pt if pt else None,
                       "wall_seconds": round(sum(c["wall_seconds"] for c in cycles), 4)}
    return out, m


def c3_steps(lib, mixed, inputs, survey, warmup, steps, only):
    """warmup untimed and steps timed mesh_adapt cycles on copies of the
    survey's cycle inputs, in cycle order; returns the timed totals."""
    dll = lib.dll
    dll.ohx_set_precision(1 if mixed else 0)

    def check(k, r, st, m):
        c = survey["cycles"][k % C3_CYCLES]
        if r != c["result"] or int(st[0]) != c["passes"] or m.counts() != c["counts"]:
            raise RuntimeError("step on cycle %d differs from the survey run: %r vs %r" % (
                k % C3_CYCLES, (r, int(st[0]), m.counts()), (c["result"], c["passes"], c["counts"])))

    for k in range(warmup):
        m = inputs[k % C3_CYCLES].clone()
        r, st = m.adapt_timed(ADAPT3[0], ADAPT3[1], ADAPT3[2], C3_MAX_OPS)
        check(k, r, st, m)
        del m
    work = [inputs[(warmup + k) % C3_CYCLES].clone() for k in range(steps)]
    dll.ohx_profile_reset()
    dll.ohx_profile_only(only.encode() if only else None)
    dll.ohx_profile_enable(1 if only else 0)
    dll.ohx_synchronize()
    elems = passes = pass_s = 0.0
    res = []
    t0 = time.perf_counter()
    for k in range(steps):
        r, st = work[k].adapt_timed(ADAPT3[0], ADAPT3[1], ADAPT3[2], C3_MAX_OPS)
        res.append((r, st))
        elems += st[1]
        passes += st[0]
        pass_s += st[2]
    dll.ohx_synchronize()
    wall = time.perf_counter() - t0
    dll.ohx_profile_enable(0)
    for k in range(steps):
        check(warmup + k, res[k][0], res[k][1], work[k])
    del work
    return {"wall_seconds": wall, "elements_over_passes": elems, "passes": passes, "pass_seconds": pass_s,
            "value": elems / wall, "per_pass_metric": elems / pass_s if pass_s else None}


def c3_cpu_baseline(lib):
    """The isotropic stand-in of c3 (v1 has only a scalar size): the 64^3 box,
    adapt_size = v1 size_from_hessian of the same analytic Hessian (9
    components, weight 1, min 0.002, max 0.1), mesh_adapt until C3_CPU_OPS
    operations.  Timed on the reference's OpenMP build (all host threads) and
    run by the product's FP64 path on the GPU; final meshes compared."""
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    ref = _checker("ref_omp")
    if ref is None:
        return None

    def standin(l, on_device):
        if on_device:
            m = S.Mesh.box_on_device(l, 3, C3_N)
        else:
            x, c = S.box(3, C3_N)
            m = S.Mesh(l, 3, x, c)
        m.derive_model_full()
        h6 = W.tanh_front_hessian(m.coords().reshape(-1, 3), a=C3_A, s0=C3_S0, dim=3)
        h9 = h6[:, [0, 1, 2, 1, 3, 4, 2, 4, 5]]
        m.add_field(0, "hessian9", h9.ravel())
        m.size_from_hessian("hessian9", None, C3_HMIN, C3_HMAX)
        m.free_field(0, "hessian9")
        return m

    t0 = time.perf_counter()
    mc = standin(ref, False)
    setup = time.perf_counter() - t0
    print("c3 cpu baseline: stand-in set up in %.1f s" % setup, file=sys.stderr, flush=True)
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)  # the reference prints per-pass lines
    try:
        r, st = mc.adapt_timed(ADAPT3[0], ADAPT3[1], ADAPT3[2], C3_CPU_OPS)
    finally:
        _LIBC.fflush(None)
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)
    print("c3 cpu baseline: %d passes in %.1f s" % (int(st[0]), st[2]), file=sys.stderr, flush=True)
    lib.dll.ohx_set_precision(0)
    mg = standin(lib, True)
    rg, sg = mg.adapt_timed(ADAPT3[0], ADAPT3[1], ADAPT3[2], C3_CPU_OPS)
    same = (r == rg) and light_digest(mc) == light_digest(mg)
    return {
        "value": st[1] / st[2] if st[2] else None,
        "unit": "elements/s",
        "cores": threads,
        "kind": "reference",
        "sample": "reference OpenMP loop build, %d threads: the first %d passes of mesh_adapt on the isotropic "
                  "stand-in of this workload (64^3 box, adapt_size from v1 size_from_hessian of the same Hessian, "
                  "min 0.002, max 0.1), %.0f input elements in %.1f s of passes (set-up %.1f s untimed); v1 "
                  "cannot express the anisotropic metric" % (threads, int(st[0]), st[1], st[2], setup),
        "gpu_fp64_same_standin": {"value": sg[1] / sg[2] if sg[2] else None, "passes": int(sg[0]),
                                  "same_final_mesh_as_cpu_reference": bool(same)},
    }


def run_c3(args, lib):
    dll = lib.dll
    dll.ohx_set_indset_max_rounds(C3_INDSET_ROUNDS)
    dll.ohx_indset_rounds_max(1)
    surveys, finals = {}, {}
    for prec, mixed in (("fp64", False), ("mixed", True)):
        sv, m = c3_survey(lib, mixed)
        surveys[prec] = sv
        finals[prec] = m
        print("c3 survey %s: %d passes, %.3g el/s per pass" % (
            prec, sv["full_run"]["passes"], sv["full_run"]["per_pass_metric"]), file=sys.stderr, flush=True)
    rounds_seen = int(dll.ohx_indset_rounds_max(1))
    # decision identity: every cycle's counts and min quality, and the
    # final mesh bit for bit
    identical_cycles = all(
        surveys["fp64"]["cycles"][k]["counts"] == surveys["mixed"]["cycles"][k]["counts"]
        and surveys["fp64"]["cycles"][k]["min_quality"] == surveys["mixed"]["cycles"][k]["min_quality"]
        for k in range(C3_CYCLES))
    final_digest = {p: digest(finals[p]) for p in finals}
    identical_final = final_digest["fp64"] == final_digest["mixed"]
    # cycle inputs are the same for both precisions (identical decisions):
    # the timed steps of both run on the FP64 survey's inputs
    inputs = surveys["fp64"]["inputs"]
    surveys["mixed"]["inputs"] = None
    del finals
    dom = surveys["mixed"]["dominant"]
    fp = c3_steps(lib, False, inputs, surveys["fp64"], args.warmup, args.steps, None)
    print("c3 timed fp64: %.3g el/s" % fp["value"], file=sys.stderr, flush=True)
    dll.ohx_mixed_stats_reset()
    mx = c3_steps(lib, True, inputs, surveys["mixed"], args.warmup, args.steps, dom)
    print("c3 timed mixed: %.3g el/s" % mx["value"], file=sys.stderr, flush=True)
    roof = roofline_of(dll, dom, "c3")
    mstats = mixed_stats(dll)
    dll.ohx_set_precision(0)
    tuning = tuning_report(dll)
    del inputs
    surveys["fp64"]["inputs"] = None
    cpu = None if args.no_cpu else c3_cpu_baseline(lib)
    dll.ohx_set_indset_max_rounds(0)
    svm, svf = surveys["mixed"], surveys["fp64"]
    out = {
        "metric": "elements/sec per adapt pass, FP32-mixed vs FP64 on 1 GPU, at equal min quality",
        "value": round(mx["value"], 1),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mx["wall_seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "c3: 3D Kuhn box 64^3x6 = 1572864 tets, metric from the Hessian of tanh(50(x+y+z-1.5)) "
                        "(FP64 eigendecomposition, eps 0.01, h in [0.002, 0.1]), 8 cycles of mesh_adapt; a step "
                        "is one mesh_adapt cycle (cycle k mod 8 of the sequence, from its input mesh)",
            "precision": "FP32-mixed: split, collapse and swap qualities in FP32 (reject-only screen, or certified "
                         "accept/reject and goodness comparisons), FP64 re-evaluation of every decision within 1e-4; "
                         "the variant per kernel and size class is the fastest measured, FP64 included",
            "max_ops": C3_MAX_OPS,
            "ops_per_cycle": [c["passes"] for c in svf["cycles"]],
            "cycles_at_op_limit": sum(1 for c in svf["cycles"] if c["result"] < 0),
            "indset_max_rounds": C3_INDSET_ROUNDS,
            "indset_rounds_needed_max": rounds_seen,
            "indset_note": "the reference stops an independent set after 100 rounds (indset.cpp:33) and would "
                           "abort this workload in its first cycle; the limit is raised here",
            "timed_passes": {"fp64": fp["passes"], "mixed": mx["passes"]},
            "per_pass_metric": {"fp64": fp["per_pass_metric"], "mixed": mx["per_pass_metric"]},
        },
        "fp64_value": round(fp["value"], 1),
        "mixed_over_fp64": mx["value"] / fp["value"],
        "decisions": {
            "identical_final_mesh": bool(identical_final),
            "identical_every_cycle_counts_and_min_quality": bool(identical_cycles),
            "final_min_quality": {"fp64": svf["final_min_quality"], "mixed": svm["final_min_quality"]},
            "equal_final_min_quality": svf["final_min_quality"] == svm["final_min_quality"],
            "mixed_stats_timed": mstats,
            "mixed_stats_survey": svm["mixed"],
        },
        "full_run": {"fp64": svf["full_run"], "mixed": svm["full_run"]},
        "per_cycle": {p: [{"passes": c["passes"], "elements": c["elements"], "pass_seconds": round(c["pass_seconds"], 4),
                           "tets_out": c["counts"][3]} for c in surveys[p]["cycles"]] for p in ("fp64", "mixed")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": svm["kernels"],
        "kernels_fp64": svf["kernels"],
        "kernels_note": "survey runs (the complete 8-cycle sequence) with every kernel timed by dispatch-carried "
                        "events; value and roofline come from the timed steps, where only the dominant kernel "
                        "carries events",
        "host_sync": svm["host_sync"],
        "host_malloc": svm["host_malloc"],
        "peak_device_bytes": max(svm["peak_device_bytes"], svf["peak_device_bytes"]),
        "host_field_seconds": round(svm["host_field_seconds"], 4),
        "autotune": tuning,
    }
    print(json.dumps(out))


# -------------------------------------------------------------------- c2
def c2_prepare(lib, on_device):
    if on_device:
        m = S.Mesh.box_on_device(lib, 2, N_CELLS)
    else:
        x, c = S.box(2, N_CELLS)
        m = S.Mesh(lib, 2, x, c)
    m.derive_model_full()
    m.add_field(0, "adapt_size", W.uniform_size(m.count(0), H))
    xs = m.coords().reshape(-1, 3)
    m.add_field(0, "warp", W.sine_warp_2d(xs).ravel())
    m.warp_to_limit(WARP_FLOOR)
    return m


def c2_cpu_baseline(product_digest, product_passes):
    """The reference itself on this host's cores: its OpenMP loop build with
    every thread OMP_NUM_THREADS allows (the reported baseline) and its
    serial build (one core), ~10 s of repeated full adapts each."""
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    out = None
    for which, cores in (("ref_omp", threads), ("ref", 1)):
        lib = _checker(which)
        if lib is None:
            continue
        total_el = total_pass_t = total_t = 0.0
        runs, parity = 0, None
        t_start = time.time()
        while runs < 1 or (time.time() - t_start < 10.0 and runs < 50):
            mm = c2_prepare(lib, False)
            devnull = os.open(os.devnull, os.O_WRONLY)
            saved = os.dup(1)
            os.dup2(devnull, 1)
            try:
                r, st = mm.adapt_timed(*ADAPT_ARGS)
            finally:
                _LIBC.fflush(None)
                os.dup2(saved, 1)
                os.close(devnull)
                os.close(saved)
            total_el += st[1]
            total_pass_t += st[2]
            total_t += st[3]
            if runs == 0:
                parity = (digest(mm) == product_digest) and (int(st[0]) == product_passes)
            runs += 1
        rec = {"value": total_el / total_t, "unit": "elements/s", "cores": cores, "kind": "reference",
               "sample": "%d full mesh_adapt runs of the same workload (%.1f s), reference %s loop build" % (
                   runs, total_t, "OpenMP" if which == "ref_omp" else "serial"),
               "per_pass_metric": total_el / total_pass_t if total_pass_t else None,
               "same_result_as_gpu": bool(parity)}
        if out is None:
            out = rec
        else:
            out.setdefault("other_builds", []).append(rec)
    return out


def run_c2(args, lib):
    dll = lib.dll
    base = c2_prepare(lib, True)
    t_in = base.count(2)
    for _ in range(args.warmup):
        m = base.clone()
        m.adapt_timed(*ADAPT_ARGS)
    dll.ohx_synchronize()
    dll.ohx_profile_reset()
    dll.ohx_profile_only(None)
    dll.ohx_profile_enable(1)
    ref_m = base.clone()
    r_ref, st_ref = ref_m.adapt_timed(*ADAPT_ARGS)
    dll.ohx_synchronize()
    dll.ohx_profile_enable(0)
    tab, host_sync, host_malloc = kernel_table(dll)
    dom = max(tab, key=lambda k: tab[k][0]) if tab else None
    ref_digest = digest(ref_m)
    inputs = [base.clone() for _ in range(args.steps)]
    dll.ohx_profile_reset()
    dll.ohx_profile_only(dom.encode() if dom else None)
    dll.ohx_profile_enable(1)
    dll.ohx_synchronize()
    elems = passes = pass_time = 0.0
    res = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r, st = inputs[k].adapt_timed(*ADAPT_ARGS)
        res.append(r)
        elems += st[1]
        passes += st[0]
        pass_time += st[2]
    dll.ohx_synchronize()
    wall = time.perf_counter() - t0
    dll.ohx_profile_enable(0)
    dll.ohx_profile_only(None)
    if any(r != r_ref for r in res) or digest(inputs[-1]) != ref_digest:
        raise RuntimeError("c2: a timed step differs from the survey step")
    roof = roofline_of(dll, dom, "c2")
    cpu = None if args.no_cpu else c2_cpu_baseline(ref_digest, int(st_ref[0]))
    out = {
        "metric": "elements/sec per adapt pass (FP64 drop-in, 1 GPU)",
        "value": round(elems / wall, 1), "unit": "elements/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c2: 2D Kuhn square 256x256x2 = %d triangles, sine warp 0.1 via warp_to_limit(0.2), "
                               "mesh_adapt h=1/200 (FP64 drop-in); a step is one mesh_adapt" % t_in,
                   "passes_per_step": passes / args.steps, "elements_per_step": elems / args.steps,
                   "per_pass_metric": round(elems / pass_time, 1) if pass_time else None},
        "roofline": roof, "cpu_baseline": cpu, "kernels": per_kernel(tab),
        "kernels_note": "one untimed step with every kernel timed by dispatch-carried events",
        "host_sync": {"count": host_sync[1], "ms": round(host_sync[0], 3)} if host_sync else None,
        "peak_device_bytes": int(dll.ohx_device_bytes_high_water()), "autotune": tuning_report(dll),
    }
    print(json.dumps(out))


# -------------------------------------------------------------------- c5
C5_N = 160
C5_OPS = 40


def run_c5(args, lib):
    """The first C5_OPS passes of the first cycle (the cycle itself needs
    thousands: each coarsen pass removes ~0.1 % of 36 M tets), then one swap
    pass on the result: per-kernel GB/s at the real shape, peak memory."""
    dll = lib.dll
    mixed = args.precision == "mixed"
    dll.ohx_set_precision(1 if mixed else 0)
    dll.ohx_set_indset_max_rounds(C3_INDSET_ROUNDS)
    dll.ohx_device_reset_high_water()
    t0 = time.perf_counter()
    m = S.Mesh.box_on_device(lib, 3, C5_N)
    m.derive_model_full()
    dll.ohx_synchronize()
    setup = time.perf_counter() - t0
    t_in0 = m.count(3)
    x = m.coords().reshape(-1, 3)
    m.add_field(0, "metric", W.rotating_metric(x, 1e-3, 20.0, dim=3).ravel())
    dll.ohx_profile_reset()
    dll.ohx_profile_only(None)
    dll.ohx_profile_enable(1)
    r, st = m.adapt_timed(ADAPT3[0], ADAPT3[1], ADAPT3[2], C5_OPS)
    counts_after = m.counts()
    ts = time.perf_counter()
    swapped = m.swap_slivers(0.3, 4)
    dll.ohx_synchronize()
    swap_s = time.perf_counter() - ts
    dll.ohx_profile_enable(0)
    tab, host_sync, host_malloc = kernel_table(dll)
    dom = max(tab, key=lambda k: tab[k][0]) if tab else None
    out = {
        "metric": "elements/sec per adapt pass (%s, 1 GPU)" % ("FP32-mixed" if mixed else "FP64"),
        "value": round(st[1] / st[2], 1) if st[2] else None, "unit": "elements/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32+f64" if mixed else "f64", "data": "synthetic",
        "config": {"workload": "c5: 3D Kuhn box 160^3x6 = %d tets, rotating metric 1e-3 x 20:1; the first %d "
                               "operations of cycle 1 of mesh_adapt, then one swap pass" % (t_in0, C5_OPS),
                   "passes": int(st[0]), "elements_over_passes": st[1], "pass_seconds": round(st[2], 4),
                   "adapt_result": r, "counts_after_passes": counts_after, "swap_changed_mesh": int(swapped),
                   "swap_seconds": round(swap_s, 4), "final_counts": m.counts(), "setup_seconds": round(setup, 4)},
        "roofline": roofline_of(dll, dom, "c5"), "kernels": per_kernel(tab),
        "host_sync": {"count": host_sync[1], "ms": round(host_sync[0], 3)} if host_sync else None,
        "peak_device_bytes": int(dll.ohx_device_bytes_high_water()),
        "mixed": mixed_stats(dll) if mixed else None, "autotune": tuning_report(dll),
    }
    dll.ohx_set_precision(0)
    dll.ohx_set_indset_max_rounds(0)
    print(json.dumps(out))


# -------------------------------------------------------------------- c4
def run_c4(args, lib):
    """256 meshes one mesh_adapt call at a time, then concurrently
    (ohx_adapt_many: worker threads, one HIP stream each)."""
    dll = lib.dll
    mixed = args.precision == "mixed"
    specs = W.batch_specs(256, seed=2024)

    def part(s):
        x, c = S.box(3, s["n"])
        m = S.Mesh(lib, 3, x * s["side"], c)
        m.derive_model_full()
        m.add_field(0, "metric", W.constant_metric(m.count(0), s["rot"], s["lengths"]).ravel())
        return m

    t0 = time.perf_counter()
    parts = [part(s) for s in specs]
    t_in = sum(p.count(3) for p in parts)
    dll.ohx_synchronize()
    t_setup = time.perf_counter() - t0
    # one at a time, FP64 (the drop-in) and in the chosen precision
    seq = {}
    singles = None
    for prec in ("fp64", "mixed") if mixed else ("fp64",):
        dll.ohx_set_precision(1 if prec == "mixed" else 0)
        clones = [p.clone() for p in parts]
        dll.ohx_synchronize()
        el = 0.0
        res = []
        t0 = time.perf_counter()
        for m in clones:
            r, st = m.adapt_timed(ADAPT3[0], ADAPT3[1], ADAPT3[2], 50)
            res.append(r)
            el += st[1]
        dll.ohx_synchronize()
        wall = time.perf_counter() - t0
        seq[prec] = {"wall_seconds": round(wall, 4), "wall_metric": el / wall, "results": res}
        if singles is None:
            singles = clones
    sample = list(range(0, 256, 16))
    ref_digests = {c: digest(singles[c]) for c in sample}
    dll.ohx_set_precision(1 if mixed else 0)
    runs = {}
    for nthreads in (4, 8, 16):
        clones = [p.clone() for p in parts]
        dll.ohx_synchronize()
        t0 = time.perf_counter()
        limited, res, st = S.Mesh.adapt_many(lib, clones, ADAPT3[0], ADAPT3[1], ADAPT3[2], 50, nthreads)
        wall = time.perf_counter() - t0
        runs[nthreads] = {
            "wall_seconds": round(wall, 4), "wall_metric": st[1] / wall, "elements_over_passes": float(st[1]),
            "components_at_op_limit": int(limited),
            "identical_return_values": bool([int(v) for v in res] == [int(v) for v in seq["fp64"]["results"]]),
            "identical_sampled_meshes_to_fp64_one_at_a_time": bool(
                all(digest(clones[c]) == ref_digests[c] for c in sample)),
        }
        del clones
    best = max(runs, key=lambda k: runs[k]["wall_metric"])
    dll.ohx_set_precision(0)
    value = runs[best]["wall_metric"]
    for p in seq:
        seq[p].pop("results")
    out = {
        "metric": "elements/sec per adapt pass (256 meshes at once, %s, 1 GPU)" % ("FP32-mixed" if mixed else "FP64"),
        "value": round(value, 1), "unit": "elements/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32+f64" if mixed else "f64", "data": "synthetic",
        "config": {"workload": "c4: 256 meshes (seed 2024), n log-uniform in [8, 32], %d tets in all, random "
                               "rotation, metric axis lengths log-uniform in [0.005, 0.1]; one mesh_adapt each "
                               "(<= 50 ops)" % t_in,
                   "setup_seconds": round(t_setup, 3), "streams": best,
                   "value_is": "sum of the meshes' pass input elements / wall time of the concurrent call"},
        "concurrent": runs, "one_at_a_time": seq,
        "concurrent_over_fp64_one_at_a_time": value / seq["fp64"]["wall_metric"],
        "identity_sample": sample, "peak_device_bytes": int(dll.ohx_device_bytes_high_water()),
    }
    print(json.dumps(out))


COPY_BYTES = 1 << 28  # 256 MiB


def calibrate_copy():
    """Five device-to-device copies of 256 MiB (one copy kernel each): the
    profiler's FETCH_SIZE / WRITE_SIZE for them against the known size give
    the counters' scale."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(COPY_BYTES)) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": COPY_BYTES, "copies": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5"], default="c3")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                    help="c4/c5 only (c3 always runs both)")
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the whole mesh lives in one GPU's HBM; no run here "
              "uses a second GPU); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    lib = S.product()
    setup_api(lib.dll)
    {"c2": run_c2, "c3": run_c3, "c4": run_c4, "c5": run_c5}[args.workload](args, lib)


if __name__ == "__main__":
    main()
