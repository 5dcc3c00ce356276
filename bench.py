"""Benchmark of the fermion hot path on one MI355X.

Workload (the single-RHS FP64 drop-in at 64^3): sigma drawn from the
reference's own random stream (dSFMT-132049 after Initrand(1231)) as
Gauss2MeanVariance(0, 2/(N_f beta_s)) pairs with N_f = 4, beta_s = 0.85 on
the V dual sites, sigma mass from it, a complex Z2 source (re and im each
Casuale() <= 0.5 ? 1 : -1), CG on M M^dag to the absolute residual 1e-10 (the
reference's exit test, conjugategradient.c:87-89).  One step = one CG solve
with inputs resident in HBM.

Reported: RHS-solves/s of the FP64 drop-in (the headline `value`), the Dirac
application rate and GB/s, mixed-precision and batched solve rates, the same
legs at the configurations' other sizes (32^3, 96^3), the batched condensate
(48^3 x 64 sources) and the uneven ensembles; the roofline of the dominant
kernels (the CG iteration pair) from HIP events on their stream, with the
HBM traffic per iteration recorded from rocprofv3 counter runs of the same
kernels; the graph-node and grid-barrier probes behind the CG schedule; and
the serial CPU baseline (the oracle, 1 core: the median of five full solves
of the same system) with the reference's own O(V^2) set-up as an estimate.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--L 64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-gross-neveu-model_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_DIRAC_F64 = 40      # 2*16 + 8 per site per M or M^dag application
BYTES_CG_ITER_F64 = 144   # 8*16 + 2*8 per site per CG iteration
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_JSON = os.path.join(PROFILES, "cg_f64_L64_traffic.json")
# sources whose change invalidates a recorded traffic measurement
KERNEL_SOURCES = ["gnx_cgm.hip", "gnx_cgstate.h", "gnx_solver.h", "gnx_stencil.h", "gnx_common.h", "gnx_engine.h",
                  "gnx_engine.hip"]


def kernel_sources_sha():
    import hashlib
    h = hashlib.sha256()
    for n in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "3d-gross-neveu-model_amd", "csrc", n), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def shape_tag(shape):
    # kernel template arguments: <sites per thread, XCD slab mapping, link-coupled operator>,
    # prefixed "cg1" for the one-reduction kernels (k_cg1_X, k_cg1_Y), "cgd" for
    # their deferred-update form (k_cgd_X, k_cgd_Y), "cgf" for the fused form (k_cgf_Z)
    pre = ("cgf" if shape.get("fused") else "cgd" if shape.get("deferred") else "cg1" if shape.get("one_reduction")
           else "")
    return pre + f"<{shape['sites_per_thread']}, {'true' if shape['xcd_slab'] else 'false'}, false>"


CG_KERNELS = ("k_cgm_A", "k_cgm_B", "k_cg1_X", "k_cg1_Y", "k_cgd_X", "k_cgd_Y")


def cg_kernel_label(shape):
    if shape and shape.get("deferred"):
        return "k_cgd_X + k_cgd_Y (one CG iteration, one-reduction form, deferred x and p updates)"
    if shape and shape.get("one_reduction"):
        return "k_cg1_X + k_cg1_Y (one CG iteration, one-reduction form)"
    return "k_cgm_A + k_cgm_B (one CG iteration)"


def recorded_traffic(L, shape):
    """HBM bytes per CG iteration measured by rocprofv3 for these kernel
    sources and launch shape (see summarize_rocprof), else None."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except OSError:
        return None
    if t.get("sources_sha256") != kernel_sources_sha() or t.get("L") != L:
        return None
    return t.get("bytes_per_iteration", {}).get(shape_tag(shape))


def summarize_rocprof(kt_dir, fetch_dir, write_dir, L, untraced=None):
    """Summaries of rocprofv3 runs of this bench (no GPU needed): the
    kernel-trace statistics of the CG kernels, and HBM bytes per CG
    iteration from FETCH_SIZE / WRITE_SIZE passes, each scaled by what the
    counter reads for the plain device copy of --copy-calibration.  With
    `untraced` (the JSON line of an unprofiled run of the same build), the
    traced durations are reconciled with the unprofiled clock: the tracer
    lengthens each dispatch, by an amount the graph-node probes measure
    (their traced duration against their unprofiled time per node), and the
    CG kernels' durations less that amount must add up to the unprofiled
    time per iteration."""
    import csv
    import statistics

    def short(n):
        return n.split("(")[0].replace("void ", "").replace("gnx::", "").replace("gnx_probe::", "")

    out = []
    stats = list(csv.DictReader(open(os.path.join(kt_dir, "run_kernel_stats.csv"))))
    out.append("| Kernel | Calls | Average (ns) | Min (ns) | Max (ns) | Percent |")
    out.append("|---|---:|---:|---:|---:|---:|")
    for r in stats[:16]:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']):.0f} | {r['MinNs']} | "
                   f"{r['MaxNs']} | {float(r['Percentage']):.2f} |")
    avg = {short(r["Name"]): float(r["AverageNs"]) / 1e3 for r in stats}
    # the probes also run at other lattice sizes (the stream_152B_<n>^3
    # nodes): their averages are taken over the dispatches at L^3 only
    probe_avg = {}
    tp = os.path.join(kt_dir, "run_kernel_trace.csv")
    if os.path.exists(tp):
        acc = {}
        for r in csv.DictReader(open(tp)):
            n = short(r["Kernel_Name"])
            if n in ("k_empty", "k_stream40", "k_stream152"):
                want = 1024 * 256 if n == "k_empty" else (L ** 3 + 255) // 256 * 256
                if int(r["Grid_Size_X"]) == want:
                    acc.setdefault(n, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        probe_avg = {n: statistics.mean(v) / 1e3 for n, v in acc.items()}
    if untraced:
        u = json.load(open(untraced))
        nodes = u["probes"]["graph_nodes"]
        probe_kernel = {"empty": "k_empty", "stream_40B": "k_stream40", "stream_152B": "k_stream152"}
        out.append("")
        out.append("Tracer inflation, from the graph-node probes (64 nodes per graph, each node reading the "
                   f"previous node's output; traced average duration of the {L}^3 dispatches against the "
                   f"unprofiled time per node of the same build's bench line, `{os.path.basename(untraced)}`):")
        out.append("")
        out.append("| Probe kernel | Traced (us) | Unprofiled per node (us) | Inflation (us) |")
        out.append("|---|---:|---:|---:|")
        infl = {}
        for name, k in probe_kernel.items():
            t = probe_avg.get(k, avg.get(k))
            if t is not None and name in nodes:
                infl[name] = t - nodes[name]["us_per_node"]
                out.append(f"| `{k}` | {t:.2f} | {nodes[name]['us_per_node']:.2f} | {infl[name]:+.2f} |")
        cg = [k for k in avg if k.startswith(CG_KERNELS)]
        xk = [k for k in cg if k.startswith(("k_cg1_X", "k_cgd_X", "k_cgm_A"))]
        yk = [k for k in cg if k.startswith(("k_cg1_Y", "k_cgd_Y", "k_cgm_B"))]
        if xk and yk and "stream_40B" in infl and "stream_152B" in infl:
            x_est = avg[xk[0]] - infl["stream_40B"]
            y_est = avg[yk[0]] - infl["stream_152B"]
            per_it = u["roofline"]["us_per_iteration"]
            frac = BYTES_CG_ITER_F64 * L ** 3 / ((x_est + y_est) * 1e-6) / 1e9 / HBM_PEAK_GBS
            out.append("")
            out.append(f"CG iteration: `{xk[0]}` {avg[xk[0]]:.2f} us traced, {x_est:.2f} us less the "
                       f"40 B/site probe's inflation; `{yk[0]}` {avg[yk[0]]:.2f} us traced, {y_est:.2f} us less "
                       f"the 152 B/site probe's inflation; sum {x_est + y_est:.2f} us against {per_it:.2f} us per "
                       f"iteration unprofiled (HIP events over whole solves, graph boundaries and start kernels "
                       f"included).  From the profile: roofline frac {frac:.3f} (144 B/site); the bench line: "
                       f"{u['roofline']['frac']:.3f}.")
    counters = {}
    for ctr, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        per = {}
        copy = []
        for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
            n = short(r["Kernel_Name"])
            v = float(r["Counter_Value"])
            if n == "__amd_rocclr_copyBuffer":
                copy.append((int(r["Grid_Size"]), v))
            elif n.startswith(CG_KERNELS) and int(r["Grid_Size"]) * int(n.split("<")[1][0]) == L ** 3:
                per.setdefault(n, []).append(v)
        # the calibration's copies: the largest transfers of the largest
        # grid (a batched leg's copy-out runs on the same grid, smaller)
        big = max(g for g, _ in copy)
        top = max(v for g, v in copy if g == big)
        copy_kib = statistics.median(v for g, v in copy if g == big and v >= 0.75 * top)
        counters[ctr] = (per, copy_kib)
    calib_kib = 256 * 1024  # --copy-calibration 256: 256 MiB read and 256 MiB written
    scale = {c: calib_kib / counters[c][1] for c in counters}
    out.append("")
    out.append(f"Copy calibration (256 MiB read + 256 MiB written): FETCH_SIZE reads "
               f"{counters['FETCH_SIZE'][1]:.0f} KiB (scale {scale['FETCH_SIZE']:.3f}), WRITE_SIZE "
               f"{counters['WRITE_SIZE'][1]:.0f} KiB (scale {scale['WRITE_SIZE']:.3f}).")
    out.append("")
    out.append("| Kernel (single RHS) | Dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM MB (scaled) |")
    out.append("|---|---:|---:|---:|---:|")
    bpi = {}
    for n in sorted(set(counters["FETCH_SIZE"][0]) & set(counters["WRITE_SIZE"][0])):
        fk = statistics.median(counters["FETCH_SIZE"][0][n])
        wk = statistics.median(counters["WRITE_SIZE"][0][n])
        mb = (fk * scale["FETCH_SIZE"] + wk * scale["WRITE_SIZE"]) * 1024 / 1e6
        out.append(f"| `{n}` | {len(counters['FETCH_SIZE'][0][n])} | {fk:.0f} | {wk:.0f} | {mb:.2f} |")
        # <sites per thread, XCD slab, link-coupled, per-wave head>: the head
        # form follows from the block count, so the launch shape is the key
        tag = (("cgd" if n.startswith("k_cgd") else "cg1" if n.startswith("k_cg1") else "")
               + "<" + ", ".join(n[n.index("<") + 1:n.rindex(">")].split(", ")[:3]) + ">")
        bpi[tag] = bpi.get(tag, 0.0) + mb * 1e6
    alg = BYTES_CG_ITER_F64 * L ** 3
    out.append("")
    for tag, b in sorted(bpi.items()):
        out.append(f"Shape {tag}: {b / 1e6:.2f} MB per CG iteration (both kernels) = {b / alg:.2f} x the "
                   f"{alg / 1e6:.2f} MB algorithmic ({BYTES_CG_ITER_F64} B/site).")
    rec = {"L": L, "sources_sha256": kernel_sources_sha(), "bytes_per_iteration": bpi,
           "fetch_scale": scale["FETCH_SIZE"], "write_scale": scale["WRITE_SIZE"]}
    return "\n".join(out), rec


def make_inputs(L, seed=1231, beta_s=0.85, nflav=4, nsrc=1):
    """The reference's own random stream (dSFMT-132049 after Initrand(seed),
    the library's host RNG): sigma by Gauss2MeanVariance(0, 2/(N_f beta_s))
    pairs into (x, x + V/2) (sigma.c:29-30), then nsrc complex Z2 sources, re
    and im each Casuale() <= 0.5 ? 1 : -1 (condensate.c:57-58)."""
    import gnx
    r = gnx.Rng(seed)
    V = L * L * L
    sigma = r.sigma(V, 2.0 / (nflav * beta_s))
    bs = [r.z2_complex(V) for _ in range(nsrc)]
    return sigma, (bs[0] if nsrc == 1 else np.concatenate(bs))


def host_cpu():
    """Host CPU description for the baseline: model, logical CPUs, CPUs this
    process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def setup_estimate(sizes=(32, 64, 96)):
    """The reference pays an O(V^2) table set-up before its first solve
    (coordinate.c:253-267, twice per site from lattice.c:32-33).  Timed
    here: the reference's own StartLattice at the size its macro.h ships
    (oracle/_ref/ref_setup_time, built from the reference's sources where
    they are present), and the oracle's restatement of the same search at
    three small sizes, which shows the V^2 growth; both extrapolated by V^2."""
    import statistics
    import subprocess
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    out = {"note": "excluded from the CPU baseline's solve rate; extrapolated by V^2"}
    pts = []
    for L in (12, 16, 20):
        lat = P.Lattice(L)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            lat.tables_by_search()
            ts.append(time.perf_counter() - t0)
        pts.append((lat.V, statistics.median(ts)))
    slope = float(np.polyfit(np.log([v for v, _ in pts]), np.log([t for _, t in pts]), 1)[0])
    v0, t0s = pts[-1]
    out["oracle_search"] = {"seconds": {f"{round(v ** (1 / 3))}^3": t for v, t in pts},
                            "fitted_exponent_in_V": slope,
                            "extrapolated_seconds": {f"{L}^3": t0s * (L ** 3 / v0) ** 2 for L in sizes}}
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_setup_time")
    if os.path.exists(exe):
        try:
            vol, sec = subprocess.run([exe], capture_output=True, text=True, timeout=60,
                                      check=True).stdout.split()
            vol, sec = int(vol), float(sec)
            out["reference_StartLattice"] = {
                "volume": vol, "seconds": sec, "kind": "reference",
                "extrapolated_seconds": {f"{L}^3": sec * (L ** 3 / vol) ** 2 for L in sizes}}
        except Exception as exc:  # a measurement aid: report, do not fail the bench
            out["reference_StartLattice"] = {"error": str(exc)}
    return out


def cpu_hmc(P, gpu, ntraj=20, eps=0.1, steps=2, tol_high=1e-10, tol_medium=1e-3, L=8, beta_s=0.85):
    """The oracle's serial sigma-HMC on configuration 1's workload (its own
    copy of the reference's stream from Initrand(1231), drawn in the same
    order), timed, and its Delta-H and accept sequence against the GPU's
    (hmc_leg)."""
    h = P.HMC(P.Lattice(L), 1231, beta_s=beta_s)
    dh, acc = [], []
    t_traj = 0.0
    t0 = time.perf_counter()
    for _ in range(ntraj):
        ta = time.perf_counter()
        t = h.trajectory(eps, steps, tol_high, tol_medium)
        t_traj += time.perf_counter() - ta
        h.condensate(tol_high)
        dh.append(t.dH)
        acc.append(int(t.accepted))
    dt = time.perf_counter() - t0
    out = {"trajectories_per_s": ntraj / t_traj, "with_measurement_per_s": ntraj / dt, "cores": 1,
           "kind": "port", "sample": f"the {ntraj} trajectories of configuration 1 with one measurement each"}
    if gpu and "dH" in gpu:
        out["max_abs_dH_diff_vs_gpu"] = max(abs(a - b) for a, b in zip(dh, gpu["dH"]))
        out["same_accept_decisions_as_gpu"] = acc == gpu["accepted"]
    return out


def cpu_baseline(L, sigma, b, nsolves=5, hmc_gpu=None):
    """Oracle CG (serial, 1 core, the reference's loops and order) on the
    same fields: the median of nsolves full solves to the same absolute
    tolerance, and the median of 100 single Dirac applications; with
    hmc_gpu (the GPU's configuration-1 leg), the oracle's HMC on the same
    workload."""
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    lat = P.Lattice(L)
    sm = lat.sigma_mass(sigma)
    ts, its = [], []
    for _ in range(nsolves):
        t0 = time.perf_counter()
        _, ok, st = lat.cg(sm, b, None, tol=1e-10, start="imposed")
        ts.append(time.perf_counter() - t0)
        its.append(st.last_iters)
    t_solve = statistics.median(ts)
    # the serial Dirac application (M and M^dag alternating, 100 of them)
    v = b
    ta = []
    for k in range(50):
        t0 = time.perf_counter()
        w = lat.M(sm, v)
        t1 = time.perf_counter()
        v = lat.Mdag(sm, w)
        ta += [t1 - t0, time.perf_counter() - t1]
    t_app = statistics.median(ta)
    return {
        "value": 1.0 / t_solve,
        "unit": "RHS-solves/s",
        "cores": 1,
        "kind": "port",
        "host": host_cpu(),
        "sample": f"median of {nsolves} full serial solves of the same {L}^3 system to |r| < 1e-10 "
                  f"({its[0]} iterations each; {min(ts):.2f}-{max(ts):.2f} s per solve) by the oracle, "
                  "which keeps the reference's loops and accumulation order (table set-up excluded)",
        "ms_per_iteration": t_solve / its[0] * 1e3,
        "dirac": {"apps_per_s": 1.0 / t_app, "GB_s": BYTES_DIRAC_F64 * L ** 3 / t_app / 1e9,
                  "sample": "median of 100 serial applications (M and M^dag alternating) at the same size"},
        "setup_estimate": setup_estimate(),
        **({"hmc_8": cpu_hmc(P, hmc_gpu)} if hmc_gpu is not None else {}),
    }


def probes_leg(L=64, barrier=True):
    """The measurements behind the CG schedule (lib/libgnx_probes.so): the
    cost of one node of a replayed hipGraph (an empty kernel, and streaming
    kernels of 40 and 152 B/site at L^3, each reading the previous node's
    output), and of a grid-wide barrier inside one persistent kernel at
    several block counts, with and without bytes written between barriers,
    and of a persistent iteration's halo exchange without a barrier
    (barrier=False skips both: the barrier's cooperative launches, which the
    profiler's kernel tracing does not survive, and the halo exchange, whose
    bounded spins a traced run stretches to milliseconds)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "3d-gross-neveu-model_amd", "lib", "libgnx_probes.so"))
    V = L ** 3
    out = {"graph_nodes": {}, "grid_barrier": []}
    legs = [(0, "empty", 0, V), (1, "stream_40B", 40, V), (2, "stream_152B", 152, V), (3, "read_88B", 88, V),
            (4, "write_64B", 64, V)]
    legs += [(2, f"stream_152B_{n}^3", 152, n ** 3) for n in (32, 48, 96)]
    for kind, name, bps, sites in legs:
        us = C.c_double()
        rc = lib.gnx_probe_graph_nodes(kind, C.c_long(sites), 64, 20, C.byref(us))
        if rc != 0:
            out["graph_nodes"][name] = {"error": rc}
            continue
        d = {"us_per_node": us.value, "nodes": 64}
        if bps:
            d["GB_s"] = bps * sites / (us.value * 1e-6) / 1e9
            d["sites"] = sites
        out["graph_nodes"][name] = d
    # the same bytes through 4-, 8- and 16-byte lanes: a stream (3 arrays in,
    # 1 out, 33.5 MB each) and a seven-point gather of two arrays laid out
    # like one parity of 32 FP16 vectors at 64^3 (8.4 M 4-byte elements)
    out["lane_width"] = {"stream": [], "gather": []}
    for w in (4, 8, 16):
        us = C.c_double()
        nbytes = 33554432
        rc = lib.gnx_probe_width_nodes(w, C.c_long(nbytes), 64, 10, C.byref(us))
        out["lane_width"]["stream"].append({"bytes_per_lane": w, "bytes_per_array": nbytes,
                                            "us_per_node": us.value if rc == 0 else None,
                                            "GB_s": 4 * nbytes / (us.value * 1e3) if rc == 0 else None, "rc": rc})
    for w in (4, 16, 0):  # 0: each block's tile and halo staged in LDS first
        us = C.c_double()
        elems = 8388608
        rc = lib.gnx_probe_gather_nodes(w, C.c_long(elems), 64, 10, C.byref(us))
        out["lane_width"]["gather"].append({"bytes_per_lane": w if w else 4, "lds_staged": w == 0, "elements": elems,
                                            "us_per_node": us.value if rc == 0 else None,
                                            "unique_GB_s": 12 * elems / (us.value * 1e3) if rc == 0 else None,
                                            "rc": rc})
    # the empty node against the workgroups it dispatches
    out["empty_node_by_grid"] = []
    for threads in (256, 1024):
        for blocks in (1, 256, 1024, 2048, 4096):
            us = C.c_double()
            rc = lib.gnx_probe_empty_nodes(blocks, threads, 64, 20, C.byref(us))
            out["empty_node_by_grid"].append({"threads_per_block": threads, "blocks": blocks,
                                              "us_per_node": us.value if rc == 0 else None, "rc": rc})
    if not barrier:
        return out
    # the communication of a persistent CG iteration without a grid barrier:
    # one 1024-thread block per 16x8x8 tile (tiles x 1024 sites), two
    # neighbour-face exchanges and one all-block partial gather per round
    out["halo_exchange"] = []
    forms = [("flags", lib.gnx_probe_halo, 0, "device-scope stores, wait for them, neighbour flags"),
             ("flags+release", lib.gnx_probe_halo, 1, "as flags, device-scope release fence before a flag"),
             ("flags+acquire", lib.gnx_probe_halo, 2, "as flags, device-scope acquire fence after a wait"),
             ("tagged", lib.gnx_probe_halo_tag, 0, "round-tagged 64-bit words, no flags, no partials"),
             ("tagged+gather", lib.gnx_probe_halo_tag, 1, "tagged words, every block polls every partial"),
             ("tagged+tree", lib.gnx_probe_halo_tag, 2, "tagged words, partials through 16-block group leaders")]
    for tiles in ((4, 8, 8), (2, 4, 4)):
        for name, fn, mode, what in forms:
            us, failed, stale, per_cu = C.c_double(), C.c_int(), C.c_int(), C.c_int()
            rc = fn(mode, tiles[0], tiles[1], tiles[2], 200, C.c_long(98304), C.byref(us), C.byref(failed),
                    C.byref(stale), C.byref(per_cu))
            out["halo_exchange"].append({"form": name, "what": what, "blocks": tiles[0] * tiles[1] * tiles[2],
                                         "sites": tiles[0] * tiles[1] * tiles[2] * 1024,
                                         "us_per_round": us.value if rc == 0 else None,
                                         "wait_timed_out": bool(failed.value), "stale_values": stale.value,
                                         "blocks_per_cu": per_cu.value, "rc": rc})
    for mode, poll in ((0, "acquire load per poll"), (1, "relaxed polls, one acquire fence")):
        for blocks in (256, 512, 1024):
            for per_block in (0, 32768):
                us, mx, failed = C.c_double(), C.c_int(), C.c_int()
                rc = lib.gnx_probe_grid_barrier(mode, blocks, 200, C.c_long(per_block), C.byref(us), C.byref(mx),
                                                C.byref(failed))
                out["grid_barrier"].append({"polling": poll, "blocks": blocks,
                                            "bytes_per_block_between_barriers": per_block,
                                            "us_per_barrier": us.value if rc == 0 else None,
                                            "resident_blocks_max": mx.value, "wait_timed_out": bool(failed.value),
                                            "rc": rc})
    return out


def condensate_leg(ctx_factory, torch, gnx, spec, tol):
    """Chiral condensate Tr M^-1 / V from N complex Z2 sources on one sigma
    field (condensate.c:127-145 per source): one batched solve in FP64 and in
    mixed FP32 and FP16 storage, against N single FP64 solves of the same
    sources.  Complex Z2
    has E|v|^2 = 2, so each source's estimate is Re(v^dag M^-1 v) / (2V)."""
    Lc, n = (int(v) for v in spec.split(":"))
    V = Lc ** 3
    sigma_h, noise_h = make_inputs(Lc, seed=4848, nsrc=n)
    ctx = ctx_factory(Lc)
    dev = ctx.device
    sm = ctx.field()
    ctx.sigma_mass(torch.from_numpy(sigma_h).to(dev), sm)
    noise = torch.from_numpy(noise_h).to(dev)
    X = ctx.vec(n)
    out = {"workload": f"{n} complex Z2 sources at {Lc}^3, tol {tol:g}"}
    for name, prec in (("f64", gnx.PREC_F64), ("f32", gnx.PREC_F32), ("f16", gnx.PREC_F16)):
        ctx.condensate_batch(sm, noise, n, tol, X, prec=prec)  # tune + warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, per, res = ctx.condensate_batch(sm, noise, n, tol, X, prec=prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        est = np.array(per) / 2.0
        out[f"batched_{name}"] = {"estimate": float(est.mean()),
                                  "noise_variance_of_mean": float(est.var(ddof=1) / n),
                                  "seconds": dt, "rhs_solves_per_s": n / dt,
                                  "iterations_max": max(r.iterations for r in res),
                                  "all_converged": all(r.converged for r in res)}
    # per-source FP64 reference: the same sources one at a time
    x1 = ctx.vec()
    y1 = ctx.vec()
    per1 = []
    t0 = time.perf_counter()
    for k in range(n):
        v = noise[k * 2 * V:(k + 1) * 2 * V]
        ctx.cg(sm, v, x1, tol, start=gnx.START_IMPOSED)
        ctx.dirac(gnx.OP_MDAG, sm, x1, y1)
        per1.append(ctx.dot(v, y1).real / V / 2.0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    est1 = np.array(per1)
    out["per_source_f64"] = {"estimate": float(est1.mean()),
                             "noise_variance_of_mean": float(est1.var(ddof=1) / n),
                             "seconds": dt, "rhs_solves_per_s": n / dt}
    sem = float(np.sqrt(est1.var(ddof=1) / n))
    for name in ("f64", "f32", "f16"):
        b = out[f"batched_{name}"]
        b["diff_vs_per_source"] = abs(b["estimate"] - est1.mean())
        b["diff_in_statistical_errors"] = b["diff_vs_per_source"] / sem
        b["speedup_vs_per_source_f64"] = b["rhs_solves_per_s"] / out["per_source_f64"]["rhs_solves_per_s"]
    return out


def ensemble_leg(gnx, sizes, ntraj, betas=(0.6, 0.75, 0.85, 1.0), eps=0.02, steps=10):
    """Independent chains on lattices of uneven size, links on: per chain and
    trajectory one sigma LFMN2VF trajectory and one LeapFrog2Link
    trajectory (10 steps each, beta = 1, pure gauge on), each followed by its
    Metropolis test.  Links start uniform in [0, 2 pi), sigma iid
    N(0, 2/(4 beta_s)).  The link trajectory follows the reference as
    written: solves with the live operator, force from the links
    (moleculardynamics.c:419-542).  Run twice on identical chains and draws:
    every chain with its own context, stream and host thread (kernels of
    different chains run concurrently on the GPU), then all chains in
    lockstep (gnx.Ensemble: each solve point of an integrator one
    uneven-batch launch over all chains)."""
    import concurrent.futures as cf
    def make_chains(force_prec=None, linked=False):
        out = []
        for L in sizes:
            for bs in betas:
                V = L ** 3
                seed = 5000 + 10 * L + int(bs * 100)
                rng = gnx.Rng(seed)
                ctx = gnx.Context(L)
                h = gnx.HMC(ctx, beta_s=bs)
                h.set_sigma(rng.sigma(V, 2.0 / (4 * bs)))
                if force_prec is not None:
                    h.set_force_precision(force_prec)
                lh = gnx.LinkHMC(ctx, beta=1.0, pure_gauge=True, linked=linked)
                _, sm = h.get_sigma()
                if linked:
                    # near unity: phases 2 pi (u - 1/2) / 100, u uniform
                    U = np.exp(1j * 2.0 * np.pi * (rng.casuale(3 * V) - 0.5) / 100.0)
                else:
                    U = np.exp(1j * 2.0 * np.pi * rng.casuale(3 * V))
                lh.set_fields(U, sm)
                out.append({"L": L, "beta_s": bs, "V": V, "ctx": ctx, "h": h, "lh": lh, "rng": rng,
                            "dH_sigma": [], "dH_link": [], "acc_sigma": 0, "acc_link": 0, "cg_iterations": 0})
        return out

    chains = make_chains()

    def run(c, n):
        V, rng = c["V"], c["rng"]
        for _ in range(n):
            st, t = c["h"].trajectory(rng.gauss1(V), rng.gauss1(V), rng.gauss1(V), eps, steps, 1e-10, 1e-3)
            a = bool(t.ok) and rng.casuale() <= np.exp(-t.dH)
            c["h"].accept(a)
            _, sm = c["h"].get_sigma()
            c["lh"].set_fields(None, sm)
            st2, tl = c["lh"].trajectory(rng.gauss2_pairs(3 * V), rng.gauss1(V), rng.gauss1(V), eps, steps, 1e-10)
            al = bool(tl.ok) and rng.casuale() <= np.exp(-tl.dH)
            c["lh"].accept(al)
            c["dH_sigma"].append(t.dH)
            c["dH_link"].append(tl.dH)
            c["acc_sigma"] += int(a)
            c["acc_link"] += int(al)
            c["cg_iterations"] += t.cg_iterations + tl.cg_iterations
        return n

    # one warm-up trajectory per chain (each context tunes its solver shapes)
    with cf.ThreadPoolExecutor(max_workers=len(chains)) as ex:
        list(ex.map(lambda c: run(c, 1), chains))
    for c in chains:
        c["dH_sigma"].clear(), c["dH_link"].clear()
        c["acc_sigma"] = c["acc_link"] = c["cg_iterations"] = 0
    def replay_totals():
        st = [c["ctx"].replay_stats() for c in chains]
        return sum(r["slots"] for r in st), sum(r["polls"] for r in st)
    r0 = replay_totals()
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=len(chains)) as ex:
        list(ex.map(lambda c: run(c, ntraj), chains))
    dt = time.perf_counter() - t0
    r1 = replay_totals()
    out = {"workload": f"{len(chains)} chains on {sizes}^3 x beta_s {list(betas)}, links on (beta 1), "
                       f"{ntraj} sigma + link trajectories each, {steps} steps of {eps}",
           "mode": "one context, stream and host thread per chain",
           "trajectories_per_s": len(chains) * ntraj / dt, "seconds": dt,
           "replay": {"slots_launched": r1[0] - r0[0], "polls": r1[1] - r0[1]},
           "chains": [{"L": c["L"], "beta_s": c["beta_s"], "dH_sigma": c["dH_sigma"], "dH_link": c["dH_link"],
                       "accepted_sigma": c["acc_sigma"], "accepted_link": c["acc_link"],
                       "cg_iterations": c["cg_iterations"]} for c in chains]}

    # the same chains and draws, thread per chain, with the MEDIUM-tolerance
    # force solves in mixed precision (FP16 storage, FP64 reliable updates)
    mixed = make_chains(gnx.PREC_F16)
    with cf.ThreadPoolExecutor(max_workers=len(mixed)) as ex:
        list(ex.map(lambda c: run(c, 1), mixed))
    for c in mixed:
        c["dH_sigma"].clear(), c["dH_link"].clear()
        c["acc_sigma"] = c["acc_link"] = c["cg_iterations"] = 0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=len(mixed)) as ex:
        list(ex.map(lambda c: run(c, ntraj), mixed))
    dtm = time.perf_counter() - t0
    out["thread_per_chain_f16_force_solves"] = {
        "mode": "as thread per chain, the MEDIUM-tolerance (1e-3) sigma force solves in FP16 storage with "
                "FP64 reliable updates; energies and the last force in FP64",
        "trajectories_per_s": len(mixed) * ntraj / dtm, "seconds": dtm, "speedup_vs_fp64": dt / dtm,
        "max_abs_dH_sigma_diff_vs_fp64": max(abs(a - b) for c, d in zip(chains, mixed)
                                             for a, b in zip(c["dH_sigma"], d["dH_sigma"])),
        "same_accept_decisions": all(c["acc_sigma"] == d["acc_sigma"] for c, d in zip(chains, mixed))}
    for c in mixed:
        c["lh"].close()
        c["h"].close()

    # the same chains and draws in lockstep: every solve point of both
    # integrators one uneven-batch launch over all chains
    lock = make_chains()
    es = gnx.Ensemble([c["h"] for c in lock])
    el = gnx.Ensemble([c["lh"] for c in lock])

    def run_lockstep(n, lock, es, el):
        for _ in range(n):
            dr = [(c["rng"].gauss1(c["V"]), c["rng"].gauss1(c["V"]), c["rng"].gauss1(c["V"])) for c in lock]
            _, ts = es.sigma_trajectory([d[0] for d in dr], [d[1] for d in dr], [d[2] for d in dr],
                                        eps, steps, 1e-10, 1e-3)
            for c, t in zip(lock, ts):
                a = bool(t.ok) and c["rng"].casuale() <= np.exp(-t.dH)
                c["h"].accept(a)
                _, sm = c["h"].get_sigma()
                c["lh"].set_fields(None, sm)
                c["dH_sigma"].append(t.dH)
                c["acc_sigma"] += int(a)
                c["cg_iterations"] += t.cg_iterations
            dl = [(c["rng"].gauss2_pairs(3 * c["V"]), c["rng"].gauss1(c["V"]), c["rng"].gauss1(c["V"]))
                  for c in lock]
            _, tls = el.link_trajectory([d[0] for d in dl], [d[1] for d in dl], [d[2] for d in dl],
                                        eps, steps, 1e-10)
            for c, tl in zip(lock, tls):
                al = bool(tl.ok) and c["rng"].casuale() <= np.exp(-tl.dH)
                c["lh"].accept(al)
                c["dH_link"].append(tl.dH)
                c["acc_link"] += int(al)
                c["cg_iterations"] += tl.cg_iterations

    run_lockstep(1, lock, es, el)  # warm-up, as above
    for c in lock:
        c["dH_sigma"].clear(), c["dH_link"].clear()
        c["acc_sigma"] = c["acc_link"] = c["cg_iterations"] = 0
    t0 = time.perf_counter()
    run_lockstep(ntraj, lock, es, el)
    dtl = time.perf_counter() - t0
    # the two runs tune their own CG launch shapes, so rounding differs; a
    # force solve to 1e-3 can then stop one iteration apart, which moves its
    # solution by up to the tolerance
    ds = [abs(a - b) for c, d in zip(chains, lock) for a, b in zip(c["dH_sigma"], d["dH_sigma"])]
    dl = [abs(a - b) / max(1.0, abs(a)) for c, d in zip(chains, lock) for a, b in zip(c["dH_link"], d["dH_link"])]
    out["lockstep"] = {"mode": "all chains in lockstep, one uneven-batch solve per integrator solve point",
                       "trajectories_per_s": len(lock) * ntraj / dtl, "seconds": dtl,
                       "speedup_vs_thread_per_chain": dt / dtl,
                       "max_abs_dH_sigma_diff_vs_thread_per_chain": max(ds),
                       "max_rel_dH_link_diff_vs_thread_per_chain": max(dl),
                       "same_accept_decisions": all(c["acc_sigma"] == d["acc_sigma"] and c["acc_link"] == d["acc_link"]
                                                    for c, d in zip(chains, lock))}
    es.close()
    el.close()

    # the same chains and draws in lockstep, the MEDIUM-tolerance sigma force
    # solves in mixed precision (FP16 storage, FP64 reliable updates) in the
    # uneven-batch launch
    lm = make_chains(gnx.PREC_F16)
    es = gnx.Ensemble([c["h"] for c in lm])
    el = gnx.Ensemble([c["lh"] for c in lm])
    run_lockstep(1, lm, es, el)
    for c in lm:
        c["dH_sigma"].clear(), c["dH_link"].clear()
        c["acc_sigma"] = c["acc_link"] = c["cg_iterations"] = 0
    t0 = time.perf_counter()
    run_lockstep(ntraj, lm, es, el)
    dtm2 = time.perf_counter() - t0
    out["lockstep_f16_force_solves"] = {
        "mode": "as lockstep, the MEDIUM-tolerance (1e-3) sigma force solves in FP16 storage with FP64 reliable "
                "updates (one uneven-batch mixed launch per solve point); energies and the last force in FP64",
        "trajectories_per_s": len(lm) * ntraj / dtm2, "seconds": dtm2, "speedup_vs_fp64_lockstep": dtl / dtm2,
        "max_abs_dH_sigma_diff_vs_fp64": max(abs(a - b) for c, d in zip(lock, lm)
                                             for a, b in zip(c["dH_sigma"], d["dH_sigma"])),
        "same_accept_decisions": all(c["acc_sigma"] == d["acc_sigma"] for c, d in zip(lock, lm))}
    es.close()
    el.close()
    for c in chains + lock + lm:
        c["lh"].close()
        c["h"].close()

    # the consistent link-coupled theory (linked = 1: solves and force with
    # the link-coupled operator) from links near unity, in lockstep
    lk = make_chains(linked=True)
    es = gnx.Ensemble([c["h"] for c in lk])
    el = gnx.Ensemble([c["lh"] for c in lk])
    run_lockstep(1, lk, es, el)
    for c in lk:
        c["dH_sigma"].clear(), c["dH_link"].clear()
        c["acc_sigma"] = c["acc_link"] = c["cg_iterations"] = 0
    t0 = time.perf_counter()
    run_lockstep(ntraj, lk, es, el)
    dtk = time.perf_counter() - t0
    out["lockstep_linked"] = {
        "mode": "all chains in lockstep, link trajectories with the link-coupled operator (linked = 1) from "
                "links near unity (phases uniform in +-pi/100)",
        "trajectories_per_s": len(lk) * ntraj / dtk, "seconds": dtk,
        "chains": [{"L": c["L"], "beta_s": c["beta_s"], "dH_link": c["dH_link"], "accepted_link": c["acc_link"],
                    "accepted_sigma": c["acc_sigma"], "cg_iterations": c["cg_iterations"]} for c in lk]}
    es.close()
    el.close()
    for c in lk:
        c["lh"].close()
        c["h"].close()
    return out


def uneven_leg(gnx, torch, sizes, tol, betas=(0.6, 0.75, 0.85, 1.0)):
    """Independent systems on lattices of uneven size, one per (L, beta_s):
    sigma iid N(0, 2/(4 beta_s)), a complex Z2 source, FP64 CG to tol
    (imposed start).  Solved three ways: all in one uneven batch
    (gnx_multi_cg: one graph-replayed kernel pair per iteration over every
    system), one after another on per-size contexts, and concurrently, each
    on its own context, stream and host thread."""
    import concurrent.futures as cf
    dev = torch.device("cuda", 0)
    systems, ctx_of = [], {}
    for L in sizes:
        ctx_of[L] = gnx.Context(L)
        for bs in betas:
            V = L ** 3
            sig_h, b_h = make_inputs(L, seed=7000 + 10 * L + int(bs * 100), beta_s=bs)
            sm = ctx_of[L].field()
            ctx_of[L].sigma_mass(torch.from_numpy(sig_h).to(dev), sm)
            b = torch.from_numpy(b_h).to(dev)
            systems.append({"L": L, "beta_s": bs, "V": V, "sm": sm, "b": b})
    n = len(systems)
    out = {"workload": f"{n} systems on {sizes}^3 x beta_s {list(betas)}, own sigma each, "
                       f"complex Z2 source, FP64 to tol {tol:g}"}
    # one uneven batch
    m = gnx.Multi([s["L"] for s in systems])
    xs = [torch.zeros(2 * s["V"], dtype=torch.float64, device=dev) for s in systems]
    args = ([s["sm"] for s in systems], [s["b"] for s in systems], xs, tol)
    m.cg(*args, start=gnx.START_IMPOSED)  # graph capture + warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, res = m.cg(*args, start=gnx.START_IMPOSED)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["one_batch"] = {"rhs_solves_per_s": n / dt, "seconds": dt, "all_converged": st == 0,
                        "iterations": [r.iterations for r in res]}
    # one after another
    x1 = [torch.zeros_like(x) for x in xs]
    for s, x in zip(systems, x1):  # tune + warm each size
        ctx_of[s["L"]].cg(s["sm"], s["b"], x, tol, start=gnx.START_IMPOSED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, x in zip(systems, x1):
        ctx_of[s["L"]].cg(s["sm"], s["b"], x, tol, start=gnx.START_IMPOSED)
    torch.cuda.synchronize()
    dt1 = time.perf_counter() - t0
    out["one_after_another"] = {"rhs_solves_per_s": n / dt1, "seconds": dt1}
    # concurrently: a context, stream and host thread per system
    ctxs = [gnx.Context(s["L"], stream=torch.cuda.Stream(dev)) for s in systems]
    x2 = [torch.zeros_like(x) for x in xs]

    def solve(k):
        ctxs[k].cg(systems[k]["sm"], systems[k]["b"], x2[k], tol, start=gnx.START_IMPOSED)
        ctxs[k].synchronize()
    with cf.ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(solve, range(n)))  # tune + warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(solve, range(n)))
    torch.cuda.synchronize()
    dt2 = time.perf_counter() - t0
    out["concurrent_streams"] = {"rhs_solves_per_s": n / dt2, "seconds": dt2}
    out["speedup_vs_one_after_another"] = dt1 / dt
    out["speedup_vs_concurrent_streams"] = dt2 / dt
    out["max_rel_diff_vs_single"] = max(float((a - b).norm() / b.norm()) for a, b in zip(xs, x1))
    m.close()
    for c in ctxs + list(ctx_of.values()):
        c.close()
    return out


def dirac_graph_rate(gnx, torch, L, sm, b):
    """Device-side rate of back-to-back single applications: 100 (M then
    M^dag, each reading the other's output) captured once in a HIP graph on a
    dedicated stream and replayed, so no host launch sits between them;
    every replay is checked against the eager result of the same sequence."""
    V = L ** 3
    dev = sm.device
    try:
        gs = torch.cuda.Stream(dev)
        gctx = gnx.Context(L, stream=gs)
        go1, go2 = gctx.vec(), gctx.vec()
        with torch.cuda.stream(gs):
            for _ in range(50):
                gctx.dirac(gnx.OP_M, sm, b, go1)
                gctx.dirac(gnx.OP_MDAG, sm, go1, go2)
        gs.synchronize()
        ref2 = go2.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=gs):
            for _ in range(50):
                gctx.dirac(gnx.OP_M, sm, b, go1)
                gctx.dirac(gnx.OP_MDAG, sm, go1, go2)
        same = True
        # replay() launches on the current stream: every replay and both
        # timing events go to gs
        with torch.cuda.stream(gs):
            for _ in range(3):
                go2.zero_()
                graph.replay()
                gs.synchronize()
                same = same and bool(torch.equal(go2, ref2))
            g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            g0.record(gs)
            for _ in range(10):
                graph.replay()
            g1.record(gs)
        g1.synchronize()
        t_g = g0.elapsed_time(g1) * 1e-3 / 1000
        ws = 2 * 16 * V + 8 * V  # in, out, sigma mass
        res = {"apps_per_s": 1.0 / t_g, "GB_s": BYTES_DIRAC_F64 * V / t_g / 1e9,
               "frac_hbm": BYTES_DIRAC_F64 * V / t_g / 1e9 / HBM_PEAK_GBS,
               "replays_match_eager": same, "applications_per_graph": 100,
               "note": f"working set {ws / 2**20:.1f} MiB: served from the 256 MB Infinity Cache "
                       "when it fits there, so frac_hbm can exceed 1"}
        del graph, gctx, go1, go2, ref2
        return res
    except Exception as exc:  # the capture is a measurement aid, not the product path
        return {"error": str(exc)}


def size_leg(gnx, torch, L, tol, steps, nbatch, precs):
    """Configurations 2 and 3 at a lattice size other than the headline's:
    the same inputs (the reference's stream, seed 1231, nbatch sources), the
    FP64 single-RHS drop-in (RHS-solves/s over `steps` solves, iterations,
    us per iteration, the CG iteration's algorithmic bandwidth and true
    residual), the Dirac application rate, and batched FP64 and mixed
    solves of nbatch sources with iterations, reliable updates, true
    residual and error against the FP64 solution."""
    V = L ** 3
    sigma_h, B_h = make_inputs(L, nsrc=max(1, nbatch))
    ctx = gnx.Context(L)
    dev = ctx.device
    sm = ctx.field()
    ctx.sigma_mass(torch.from_numpy(sigma_h).to(dev), sm)
    B = torch.from_numpy(B_h).to(dev)
    b = B[:2 * V]
    x = ctx.vec()
    out = {"workload": f"{L}^3, sigma and {max(1, nbatch)} complex Z2 sources from the reference's stream "
                       f"(seed 1231), tol {tol:g}"}
    ctx.cg(sm, b, x, tol, start=gnx.START_IMPOSED)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(ctx.stream)
    for _ in range(steps):
        st, res = ctx.cg(sm, b, x, tol, start=gnx.START_IMPOSED)
    e1.record(ctx.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_solve = e0.elapsed_time(e1) * 1e-3 / steps
    tr = ctx.vec()
    ctx.dirac(gnx.OP_MMDAG, sm, x, tr)
    torch.cuda.synchronize()
    f64 = steps / wall
    out["f64_single"] = {"rhs_solves_per_s": f64, "iterations": res.iterations,
                         "us_per_iteration": t_solve / res.iterations * 1e6,
                         "algorithmic_GB_s": BYTES_CG_ITER_F64 * V * res.iterations / t_solve / 1e9,
                         "frac_hbm": BYTES_CG_ITER_F64 * V * res.iterations / t_solve / 1e9 / HBM_PEAK_GBS,
                         "true_residual": float(torch.linalg.norm(b - tr).item()),
                         "shape": ctx.cg_shape(1)}
    # single Dirac applications, host launched and graph-replayed
    o = ctx.vec()
    for _ in range(10):
        ctx.dirac(gnx.OP_M, sm, b, o)
    torch.cuda.synchronize()
    e0.record(ctx.stream)
    for _ in range(100):
        ctx.dirac(gnx.OP_M, sm, b, o)
        ctx.dirac(gnx.OP_MDAG, sm, o, x)
    e1.record(ctx.stream)
    e1.synchronize()
    ta = e0.elapsed_time(e1) * 1e-3 / 200
    out["dirac"] = {"apps_per_s": 1.0 / ta, "GB_s": BYTES_DIRAC_F64 * V / ta / 1e9}
    out["dirac_graph"] = dirac_graph_rate(gnx, torch, L, sm, b)
    if nbatch > 1:
        X64 = ctx.vec(nbatch)
        ctx.cg_batch(sm, B, X64, nbatch, tol, start=gnx.START_IMPOSED)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, res = ctx.cg_batch(sm, B, X64, nbatch, tol, start=gnx.START_IMPOSED)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[f"f64_x{nbatch}"] = {"rhs_solves_per_s": nbatch / dt, "iterations_max": max(r.iterations for r in res),
                                 "all_converged": all(r.converged for r in res)}
        X = ctx.vec(nbatch)
        codes = {"f32": gnx.PREC_F32, "bf16": gnx.PREC_BF16, "f16": gnx.PREC_F16}
        for pname in precs:
            ctx.cg_batch(sm, B, X, nbatch, tol, start=gnx.START_IMPOSED, prec=codes[pname])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st, res = ctx.cg_batch(sm, B, X, nbatch, tol, start=gnx.START_IMPOSED, prec=codes[pname])
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            ctx.dirac(gnx.OP_MMDAG, sm, X[:2 * V], tr)
            torch.cuda.synchronize()
            errs = [float((torch.linalg.norm(X[2 * V * k:2 * V * (k + 1)] - X64[2 * V * k:2 * V * (k + 1)])
                           / torch.linalg.norm(X64[2 * V * k:2 * V * (k + 1)])).item()) for k in range(nbatch)]
            out[f"{pname}_x{nbatch}"] = {
                "rhs_solves_per_s": nbatch / dt, "iterations_max": max(r.iterations for r in res),
                "reliable_updates_max": max(r.reliable_updates for r in res),
                "all_converged": all(r.converged for r in res),
                "true_residual_rhs0": float(torch.linalg.norm(b - tr).item()),
                "max_rel_err_vs_f64": max(errs),
                "speedup_vs_f64_single": nbatch / dt / f64}
        del X, X64
    ctx.close()
    return out


def hmc_leg(gnx, ntraj=20, eps=0.1, steps=2, tol_high=1e-10, tol_medium=1e-3, L=8, beta_s=0.85):
    """Configuration 1: sigma-only HMC at 8^3 -- LFMN2VF trajectories of 2
    steps of 0.1 (tau 0.2), START_X0_NULL force solves to 1e-3 and the
    energies' solves to 1e-10, each trajectory followed by its Metropolis
    test and one measurement (the condensate's 11 outputs) -- with every draw
    from the reference's stream (Initrand(1231): sigma, then per trajectory
    momenta, chi_0, chi_1, the Metropolis uniform, 6 V Z2 draws).  Two
    warm-up trajectories on other draws first capture this context's solver
    graphs.  Reports trajectories/s with and without the measurement, the
    Delta-H and accept sequence, and CG iterations per trajectory."""
    V = L ** 3
    ctx = gnx.Context(L)
    h = gnx.HMC(ctx, beta_s=beta_s)
    var = 2.0 / (4 * beta_s)
    w = gnx.Rng(4242)
    h.set_sigma(w.sigma(V, var))
    for _ in range(2):
        h.trajectory(w.gauss1(V), w.gauss1(V), w.gauss1(V), eps, steps, tol_high, tol_medium)
        h.accept(True)
        h.condensate(w.z2_real(6 * V), tol_high)
    rng = gnx.Rng(1231)
    h.set_sigma(rng.sigma(V, var))
    dh, acc, its, cond = [], [], [], []
    t_traj = 0.0
    t0 = time.perf_counter()
    for _ in range(ntraj):
        ta = time.perf_counter()
        st, r = h.trajectory(rng.gauss1(V), rng.gauss1(V), rng.gauss1(V), eps, steps, tol_high, tol_medium)
        a = bool(r.ok) and rng.casuale() <= np.exp(-r.dH)
        h.accept(a)
        t_traj += time.perf_counter() - ta
        cond.append(float(h.condensate(rng.z2_real(6 * V), tol_high)[0]))
        dh.append(r.dH)
        acc.append(int(a))
        its.append(r.cg_iterations)
    dt = time.perf_counter() - t0
    sig, _ = h.get_sigma()
    h.close()
    ctx.close()
    return {"workload": f"{L}^3 sigma-HMC (LFMN2VF), {ntraj} trajectories of {steps} steps of {eps}, tol "
                        f"{tol_high:g} / {tol_medium:g}, beta_s {beta_s}, one measurement per trajectory, "
                        f"draws from Initrand(1231)",
            "trajectories_per_s": ntraj / t_traj, "with_measurement_per_s": ntraj / dt,
            "acceptance": sum(acc) / ntraj, "dH": dh, "accepted": acc, "cg_iterations_per_trajectory": its,
            "condensate_tr1": cond, "sigma_mean_final": float(np.mean(sig))}


def schedule_leg(gnx, torch, L, tol, batches, reps=3):
    """The FP64 CG iteration in each kernel schedule, on the headline's
    inputs: the fixed rule's choice, the two-kernel one-reduction form
    (pinned shape 9: 2 sites per thread, slab mapping), its deferred-update
    form (15), and the fused one-kernel-per-iteration form on 4 x 8 x 16
    tiles with the halo recomputed into LDS (18, and 19 deferred).  Per
    schedule and batch: us per iteration (HIP events over `reps` solves on
    the solver's stream), iterations, and the largest relative difference
    of the solutions from the rule's."""
    V = L ** 3
    out = {"workload": f"{L}^3, sigma and complex Z2 sources from the reference's stream (seed 1231), tol {tol:g}"}
    if L % 16 or L % 8:
        return out
    saved = os.environ.pop("GNX_CG_SHAPE", None)
    try:
        for nb in batches:
            sigma_h, B_h = make_inputs(L, nsrc=nb)
            row, ref = {}, None
            for name, pin in (("rule", None), ("two_kernel_9", "9"), ("two_kernel_deferred_15", "15"),
                              ("fused_18", "18"), ("fused_deferred_19", "19")):
                if pin is None:
                    os.environ.pop("GNX_CG_SHAPE", None)
                else:
                    os.environ["GNX_CG_SHAPE"] = pin
                ctx = gnx.Context(L)
                sm = ctx.field()
                ctx.sigma_mass(torch.from_numpy(sigma_h).to(ctx.device), sm)
                B = torch.from_numpy(B_h).to(ctx.device)
                X = ctx.vec(nb)

                def solve():
                    if nb == 1:
                        st, res = ctx.cg(sm, B, X, tol, start=gnx.START_IMPOSED)
                        return [res]
                    st, res = ctx.cg_batch(sm, B, X, nb, tol, start=gnx.START_IMPOSED)
                    return list(res)

                solve()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(ctx.stream)
                for _ in range(reps):
                    res = solve()
                e1.record(ctx.stream)
                e1.synchronize()
                t = e0.elapsed_time(e1) * 1e-3 / reps
                it = max(r.iterations for r in res)
                xh = X.cpu()
                if ref is None:
                    ref = xh
                diff = max(float((torch.linalg.norm(xh[2 * V * k:2 * V * (k + 1)] - ref[2 * V * k:2 * V * (k + 1)])
                                  / torch.linalg.norm(ref[2 * V * k:2 * V * (k + 1)])).item()) for k in range(nb))
                row[name] = {"us_per_iteration": t / it * 1e6, "iterations": it, "rhs_solves_per_s": nb / t,
                             "max_rel_diff_vs_rule": diff, "shape": ctx.cg_shape(nb)}
                ctx.close()
            out[f"x{nb}"] = row
    finally:
        if saved is None:
            os.environ.pop("GNX_CG_SHAPE", None)
        else:
            os.environ["GNX_CG_SHAPE"] = saved
    return out


def small_leg(gnx, torch, sizes, tol, batches=(1, 2), reps=20):
    """Small lattices (configuration 1's 8^3 and its neighbours): the FP64
    solve in one launch per solve (the fixed rule's choice up to 12^3: one
    workgroup per system iterating to the exit) against the graph-replayed
    forms pinned by shape: 1, the two-reduction kernels (the same
    recurrences), and 7, the one-reduction kernels (the rule's choice at
    these sizes before the one-launch solve).  Per size, batch and form: us
    per iteration and per solve (HIP events over `reps` solves on the
    solver's stream), iterations, and the largest relative difference of the
    solutions from the one-launch ones."""
    out = {"workload": "L^3, sigma and complex Z2 sources from the reference's stream (seed 1231), "
                       f"imposed start, tol {tol:g}"}
    saved = os.environ.pop("GNX_CG_SHAPE", None)
    try:
        for L in sizes:
            V = L ** 3
            for nb in batches:
                sigma_h, B_h = make_inputs(L, nsrc=nb)
                row, ref = {}, None
                for name, pin in (("one_launch", None), ("graph_two_reduction_1", "1"),
                                  ("graph_one_reduction_7", "7")):
                    if pin is None:
                        os.environ.pop("GNX_CG_SHAPE", None)
                    else:
                        os.environ["GNX_CG_SHAPE"] = pin
                    ctx = gnx.Context(L)
                    sm = ctx.field()
                    ctx.sigma_mass(torch.from_numpy(sigma_h).to(ctx.device), sm)
                    B = torch.from_numpy(B_h).to(ctx.device)
                    X = ctx.vec(nb)

                    def solve():
                        st, res = ctx.cg_batch(sm, B, X, nb, tol, start=gnx.START_IMPOSED)
                        assert st == 0
                        return list(res)

                    solve()
                    torch.cuda.synchronize()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(ctx.stream)
                    for _ in range(reps):
                        res = solve()
                    e1.record(ctx.stream)
                    e1.synchronize()
                    t = e0.elapsed_time(e1) * 1e-3 / reps
                    it = max(r.iterations for r in res)
                    xh = X.cpu()
                    if ref is None:
                        ref = xh
                    diff = max(float((torch.linalg.norm(xh[2 * V * k:2 * V * (k + 1)] - ref[2 * V * k:2 * V * (k + 1)])
                                      / torch.linalg.norm(ref[2 * V * k:2 * V * (k + 1)])).item())
                               for k in range(nb))
                    row[name] = {"us_per_iteration": t / it * 1e6, "us_per_solve": t * 1e6, "iterations": it,
                                 "max_rel_diff_vs_one_launch": diff, "shape": ctx.cg_shape(nb)}
                    ctx.close()
                out[f"{L}^3 x{nb}"] = row
    finally:
        if saved is None:
            os.environ.pop("GNX_CG_SHAPE", None)
        else:
            os.environ["GNX_CG_SHAPE"] = saved
    return out


def main():
    t_start = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", default="16,32", help="batch sizes of the batched FP64 CG leg ('' to skip)")
    ap.add_argument("--mixed", default="f32,bf16,f16", help="precisions of the mixed-precision legs ('' to skip)")
    ap.add_argument("--condensate", default="48:64", metavar="L:N",
                    help="batched stochastic condensate leg: lattice L^3, N complex Z2 sources ('' to skip)")
    ap.add_argument("--chains", default="16,24,32,48", metavar="L,L,...",
                    help="uneven-ensemble leg: 4 chains per lattice size, sigma and link HMC ('' to skip)")
    ap.add_argument("--chain-trajectories", type=int, default=2)
    ap.add_argument("--sizes", default="32:32,96:16", metavar="L:N,...",
                    help="configurations 2 and 3 at other sizes: lattice L^3, FP64 single and N-source "
                         "batched FP64 and mixed solves ('' to skip)")
    ap.add_argument("--small", default="4,8,12", metavar="L,...",
                    help="small lattices: the one-launch FP64 solve against the graph-replayed forms ('' to skip)")
    ap.add_argument("--no-hmc", dest="hmc", action="store_false",
                    help="skip configuration 1 (sigma-HMC at 8^3, 20 trajectories with measurements)")
    ap.add_argument("--schedules", default="1,16", metavar="N,...",
                    help="batch sizes of the kernel-schedule comparison at the headline size ('' to skip)")
    ap.add_argument("--probes", default="all", choices=["all", "nodes", "none"],
                    help="graph-node, grid-barrier and halo-exchange probes (nodes: graph-node probes only)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also time a plain device copy of MIB MiB (the reference kernel for "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE scaling)")
    ap.add_argument("--summarize-rocprof", nargs=3, metavar=("KT_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="print the summary of rocprofv3 csv runs of this bench and record the "
                         "measured traffic for later runs (no GPU needed)")
    ap.add_argument("--untraced", metavar="BENCH_JSON",
                    help="with --summarize-rocprof: an unprofiled bench line of the same build, to reconcile "
                         "the traced durations with")
    args = ap.parse_args()
    if args.summarize_rocprof:
        text, rec = summarize_rocprof(*args.summarize_rocprof, L=args.L, untraced=args.untraced)
        print(text)
        with open(TRAFFIC_JSON, "w") as f:
            json.dump(rec, f, indent=1)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design, not a measured path", file=sys.stderr)
        return 2

    import torch
    import gnx

    L = args.L
    V = L ** 3
    sigma_h, b_h = make_inputs(L)
    ctx = gnx.Context(L)
    dev = ctx.device
    stream = ctx.stream
    sigma = torch.from_numpy(sigma_h).to(dev)
    sm = ctx.field()
    ctx.sigma_mass(sigma, sm)
    b = torch.from_numpy(b_h).to(dev)
    x = ctx.vec()
    torch.cuda.synchronize()

    # ---------------- Dirac applications ----------------
    out = ctx.vec()
    for _ in range(20):
        ctx.dirac(gnx.OP_M, sm, b, out)
    torch.cuda.synchronize()
    napp = 400
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(napp // 2):
        ctx.dirac(gnx.OP_M, sm, b, out)
        ctx.dirac(gnx.OP_MDAG, sm, out, x)
    e1.record(stream)
    e1.synchronize()
    t_app = e0.elapsed_time(e1) * 1e-3 / napp
    dirac_gbs = BYTES_DIRAC_F64 * V / t_app / 1e9
    dirac_graph = dirac_graph_rate(gnx, torch, L, sm, b)
    torch.cuda.synchronize()
    # the normal operators: M M^dag (the CG's operator) and M^dag M (offered
    # beside it), one launch with the intermediate in LDS at this size
    # (40 B/site algorithmic), against the two single-operator launches
    # through a staged vector (80 B/site)
    prod = {}
    for name, op in (("M_Mdag", gnx.OP_MMDAG), ("Mdag_M", gnx.OP_MDAGM)):
        for _ in range(10):
            ctx.dirac(op, sm, b, out)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(100):
            ctx.dirac(op, sm, b, out)
        e1.record(stream)
        e1.synchronize()
        tp = e0.elapsed_time(e1) * 1e-3 / 100
        prod[name] = {"apps_per_s": 1.0 / tp, "algorithmic_GB_s": BYTES_DIRAC_F64 * V / tp / 1e9,
                      "launches": 1 if L % 16 == 0 and L ** 3 <= 300000 else 2}
    tmp = ctx.vec()
    for _ in range(10):
        ctx.dirac(gnx.OP_MDAG, sm, b, tmp)
        ctx.dirac(gnx.OP_M, sm, tmp, out)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(100):
        ctx.dirac(gnx.OP_MDAG, sm, b, tmp)
        ctx.dirac(gnx.OP_M, sm, tmp, out)
    e1.record(stream)
    e1.synchronize()
    tp = e0.elapsed_time(e1) * 1e-3 / 100
    prod["M_Mdag_two_launches"] = {"apps_per_s": 1.0 / tp, "GB_s": 2 * BYTES_DIRAC_F64 * V / tp / 1e9,
                                   "launches": 2}
    del tmp
    # batched: 16 vectors per launch, one sigma mass (2 s_v + s_m / n per site)
    nvec = 16
    vb = torch.from_numpy(make_inputs(L, seed=77, nsrc=nvec)[1]).to(dev)
    ob = ctx.vec(nvec)
    wb = ctx.vec(nvec)
    for _ in range(4):
        ctx.dirac_batch(gnx.OP_M, sm, vb, ob, nvec)
    torch.cuda.synchronize()
    nb_launch = 100
    e0.record(stream)
    for _ in range(nb_launch // 2):
        ctx.dirac_batch(gnx.OP_M, sm, vb, ob, nvec)
        ctx.dirac_batch(gnx.OP_MDAG, sm, ob, wb, nvec)
    e1.record(stream)
    e1.synchronize()
    t_b = e0.elapsed_time(e1) * 1e-3 / (nb_launch * nvec)
    bytes_b = 32 + 8.0 / nvec
    dirac_batched = {"vectors_per_launch": nvec, "apps_per_s": 1.0 / t_b, "bytes_per_site": bytes_b,
                     "GB_s": bytes_b * V / t_b / 1e9, "frac_hbm": bytes_b * V / t_b / 1e9 / HBM_PEAK_GBS}
    del vb, ob, wb

    # ---------------- FP64 CG drop-in ----------------
    for _ in range(args.warmup):
        ctx.cg(sm, b, x, args.tol, start=gnx.START_IMPOSED)
    torch.cuda.synchronize()
    iters = []
    kt = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k0.record(stream)
        st, res = ctx.cg(sm, b, x, args.tol, start=gnx.START_IMPOSED)
        k1.record(stream)
        iters.append(res.iterations)
        kt.append((k0, k1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ktime = [a.elapsed_time(c) * 1e-3 for a, c in kt]
    solves_per_s = args.steps / wall
    it = int(np.median(iters))
    avg_k = float(np.mean(ktime))
    achieved = (BYTES_CG_ITER_F64 * V * it) / avg_k / 1e9
    # true residual of the last solution
    tr = ctx.vec()
    ctx.dirac(gnx.OP_MMDAG, sm, x, tr)
    torch.cuda.synchronize()
    true_res = float(torch.linalg.norm(b - tr).item())

    result = {
        "metric": "CG RHS-solves/s & Dirac GB/s (%HBM peak) at 64^3, FP64 vs mixed/batched @1e-10",
        "value": solves_per_s,
        "unit": "RHS-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: the reference's random stream (dSFMT-132049, Initrand(1231)): sigma by "
                "Gauss2MeanVariance(0, 2/(4*0.85)) pairs, complex Z2 source by Casuale() <= 0.5",
        "config": {"workload": f"single-RHS FP64 CG on M M^dag to |r|<{args.tol:g} at {L}^3",
                   "L": L, "V": V, "tol": args.tol, "start": "imposed", "cg_iterations": it,
                   "true_residual": true_res,
                   "relative_residual": true_res / float(torch.linalg.norm(b).item()),
                   "cg_shape": ctx.cg_shape(1)},
        # CG RHS-iterations/s: right-hand sides times CG iterations per second
        "rhs_iterations_per_s": solves_per_s * it,
        "dirac": {"apps_per_s": 1.0 / t_app, "GB_s": dirac_gbs, "frac_hbm": dirac_gbs / HBM_PEAK_GBS,
                  "bytes_per_site": BYTES_DIRAC_F64},
        "dirac_graph": dirac_graph,
        "dirac_batched": dirac_batched,
        "normal_operators": prod,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": recorded_traffic(L, ctx.cg_shape(1)),
                     "kernel": cg_kernel_label(ctx.cg_shape(1)),
                     "us_per_iteration": avg_k / it * 1e6, "solve_ms": avg_k * 1e3,
                     "bytes_per_site_iter": BYTES_CG_ITER_F64},
    }
    # ---------------- PCIe-inclusive drop-in rate ----------------
    # the reference-named drop-in receives host vectors: sigma mass, b and
    # x_in go up and x comes back on every call (pageable host memory)
    xh = np.zeros(2 * V)
    smh = sm.cpu().numpy()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        smd = torch.from_numpy(smh).to(dev)
        bd = torch.from_numpy(b_h).to(dev)
        xd = torch.from_numpy(xh).to(dev)
        ctx.cg(smd, bd, xd, args.tol, start=gnx.START_IMPOSED)
        xh = xd.cpu().numpy()
    pcie_rate = reps / (time.perf_counter() - t0)
    result["pcie_inclusive"] = {"rhs_solves_per_s": pcie_rate,
                                "note": "host sigma mass, b, x_in uploaded and x downloaded per solve"}

    # ---------------- batched FP64 CG ----------------
    batched = {}
    for nb in [int(v) for v in args.batch.split(",") if v]:
        B = torch.from_numpy(make_inputs(L, seed=7 + nb, nsrc=nb)[1]).to(dev)
        X = ctx.vec(nb)
        ctx.cg_batch(sm, B, X, nb, args.tol, start=gnx.START_IMPOSED)
        torch.cuda.synchronize()
        reps = max(1, 8 // nb)
        t0 = time.perf_counter()
        for _ in range(reps):
            st, res = ctx.cg_batch(sm, B, X, nb, args.tol, start=gnx.START_IMPOSED)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        its = [r.iterations for r in res]
        batched[str(nb)] = {"rhs_solves_per_s": nb * reps / dt, "rhs_iterations_per_s": sum(its) * reps / dt,
                            "ms_per_batch": dt / reps * 1e3,
                            "iterations_max": max(its), "iterations_min": min(its),
                            "all_converged": all(r.converged for r in res), "shape": ctx.cg_shape(nb),
                            "algorithmic_GB_s": BYTES_CG_ITER_F64 * V * sum(its) / (dt / reps) / 1e9}
        del B, X
    result["batched_f64"] = batched

    # ---------------- mixed precision (FP64 reliable updates) ----------------
    mixed = {}
    precs = {"f32": gnx.PREC_F32, "bf16": gnx.PREC_BF16, "f16": gnx.PREC_F16}
    for pname in [v for v in args.mixed.split(",") if v]:
        pcode = precs[pname]
        for nb in [1] + [int(v) for v in args.batch.split(",") if v]:
            B = b if nb == 1 else torch.from_numpy(make_inputs(L, seed=11 + nb, nsrc=nb)[1]).to(dev)
            X = ctx.vec(nb)
            ctx.cg_batch(sm, B, X, nb, args.tol, start=gnx.START_IMPOSED, prec=pcode)
            torch.cuda.synchronize()
            reps = max(2, 8 // nb)
            t0 = time.perf_counter()
            for _ in range(reps):
                st, res = ctx.cg_batch(sm, B, X, nb, args.tol, start=gnx.START_IMPOSED, prec=pcode)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            # error against the FP64 GPU solution of the same system (first RHS)
            xr = ctx.vec()
            ctx.cg(sm, B[:2 * V], xr, args.tol, start=gnx.START_IMPOSED)
            tr = ctx.vec()
            ctx.dirac(gnx.OP_MMDAG, sm, X[:2 * V], tr)
            torch.cuda.synchronize()
            mixed[f"{pname}_x{nb}"] = {
                "rhs_solves_per_s": nb * reps / dt,
                "rhs_iterations_per_s": sum(r.iterations for r in res) * reps / dt,
                "ms_per_batch": dt / reps * 1e3,
                "iterations_max": max(r.iterations for r in res),
                "reliable_updates_max": max(r.reliable_updates for r in res),
                "all_converged": all(r.converged for r in res),
                "true_residual_rhs0": float(torch.linalg.norm(B[:2 * V] - tr).item()),
                "rel_err_vs_f64_rhs0": float((torch.linalg.norm(X[:2 * V] - xr) / torch.linalg.norm(xr)).item()),
                "speedup_vs_f64_single": nb * reps / dt / solves_per_s,
                "shape": ctx.cg_mixed_shape(nb, pcode)}
            del X
    result["mixed"] = mixed
    if mixed:
        best = max(mixed.values(), key=lambda d: d["rhs_solves_per_s"])
        result["best_mixed_batched_rhs_solves_per_s"] = best["rhs_solves_per_s"]
        result["best_mixed_batched_speedup_vs_f64"] = best["speedup_vs_f64_single"]
    if args.condensate:
        result["condensate_batched"] = condensate_leg(ctx_factory=gnx.Context, torch=torch, gnx=gnx,
                                                      spec=args.condensate, tol=args.tol)
    if args.chains:
        result["ensembles_with_links"] = ensemble_leg(gnx=gnx, sizes=[int(v) for v in args.chains.split(",")],
                                                      ntraj=args.chain_trajectories)
        result["uneven_batch"] = uneven_leg(gnx=gnx, torch=torch, sizes=[int(v) for v in args.chains.split(",")],
                                            tol=args.tol)
    if args.sizes:
        result["sizes"] = {}
        for spec in args.sizes.split(","):
            Ls, n = (int(v) for v in spec.split(":"))
            result["sizes"][f"{Ls}^3"] = size_leg(gnx, torch, Ls, args.tol, args.steps, n,
                                                  [v for v in args.mixed.split(",") if v])
    if args.hmc:
        result["hmc_8"] = hmc_leg(gnx)
    if args.small:
        result["small"] = small_leg(gnx, torch, [int(v) for v in args.small.split(",")], args.tol)
    if args.schedules:
        result["schedules"] = schedule_leg(gnx, torch, L, args.tol, [int(v) for v in args.schedules.split(",")])
        # 48^3 single: the size where the rule runs the fused form (216 tiles, one per CU)
        result["schedules_48"] = schedule_leg(gnx, torch, 48, args.tol, [1])
    if args.probes != "none":
        result["probes"] = probes_leg(L, barrier=args.probes == "all")
    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        for _ in range(5):
            dst.copy_(src)
        c1.record()
        c1.synchronize()
        t = c0.elapsed_time(c1) * 1e-3 / 5
        result["copy_calibration"] = {"bytes_read": 8 * n, "bytes_written": 8 * n, "copies": 5,
                                      "GB_s": 16 * n / t / 1e9}
        del src, dst
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(L, sigma_h, b_h, hmc_gpu=result.get("hmc_8"))
    result["bench_seconds"] = time.perf_counter() - t_start  # the whole script, every leg included
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
