"""Benchmark: k-clique counting (k = 4 and 5) on a seeded R-MAT graph.

Workload (one "step"): the 4-clique and 5-clique counts of R-MAT scale 20,
edge factor 16 (a, b, c, d) = (.57, .19, .19, .05), symmetrised,
deduplicated, degree-oriented; the oriented CSR is resident in HBM before
timing starts.  Both counts come from one fused pass over the roots
(gpm_clique_all); the two single-k passes are timed beside it for reference.
value = |E+| / step time: oriented edges per second of the fused pass, which
visits every oriented edge once as the first edge of its cliques and yields
both counts (the reference's edge-rooted CFSolver runs one pass per k; the
2|E+| figure is reported beside it).

Also reported: event-timed per-kernel split, the roofline of the dominant
kernel (algorithmic bytes of the reference's edge-rooted plan for the roots
that kernel serves, over its mean duration), the measured HBM traffic of that
kernel from the committed profiler summary, a CPU baseline (the oracle timed
on a bounded sample of roots on this host, whose counts must equal the
engine's on the same roots) and extra workloads (triangle counting, learned
vs plain scheduling, labeled matching).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpm-artifact_amd"))
sys.path.insert(0, ROOT)

import gpm_amd as G  # noqa: E402

SEED = 20240520
HBM_PEAK_GBS = 8000.0
METRIC = "TC/k-clique edges/sec, matches/sec; learned-vs-plain sched speedup, exact"


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "clique_k4k5_rmat20_hbm_traffic.md")
LDS_PROFILE = os.path.join(ROOT, "profiles", "clique_k4k5_rmat20_lds.md")
# LDS peak of one MI355X: 256 CUs x 128 bytes per clock per CU (the LDS width
# CDNA documents: 32 banks x 4 bytes per clock) x 2.4 GHz peak engine clock.
LDS_PEAK_GBS = 256 * 128 * 2.4
# VALU issue peak: 256 CUs x 4 SIMDs, each issuing one wave64 VALU instruction
# per 4 clocks (16 lanes per SIMD), at 2.4 GHz -- in G instructions/s.
VALU_PEAK_GINSTR = 1024 * 2.4 / 4


def _profile_json(path, tag, kernel, scale):
    """Per-launch figure of `kernel` from a committed PMC summary line
    '<tag>: {json}', for the scale-20 workload the summaries were taken on."""
    if scale != 20 or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(tag + ":"):
                return json.loads(line.split(":", 1)[1]).get(kernel)
    return None


def measured_traffic(kernel, scale):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    return _profile_json(TRAFFIC_PROFILE, "traffic_json", kernel, scale)


def measured_lds(kernel, scale):
    """{"bytes": LDS bytes loaded + stored per launch (SQ_INSTS_LDS_LOAD_BANDWIDTH +
    SQ_INSTS_LDS_STORE_BANDWIDTH, 64-byte units), "active": LDS-active share of
    CU cycles (SQ_LDS_IDX_ACTIVE / (GRBM_GUI_ACTIVE x CUs)), "conflict": bank-conflict
    share of LDS-active cycles, "valu_instr": VALU wave64 instructions issued
    (SQ_INSTS_VALU), "valu_active": VALU-busy share of SIMD cycles} per launch."""
    return _profile_json(LDS_PROFILE, "lds_json", kernel, scale)


def roofline(dom, dom_ms, alg_bytes, traffic, lds):
    """Roofline of the dominant k-clique kernel.  Its rows live in LDS and its
    inner loops are row-word ANDs and popcounts; the PMC summary shows it
    issuing VALU instructions at ~93 % of the chip's rate with LDS at ~15 % of
    its bandwidth, so the bound is VALU issue: achieved = VALU wave64
    instructions per launch (SQ_INSTS_VALU) / the launch's event-timed
    duration, peak = 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction.
    LDS bandwidth and HBM traffic are reported beside it, and the canonical
    edge-rooted plan's bytes over the same time as a plain rate (the bitmap
    rows read each list once per root, not once per set operation, so that
    rate exceeds any memory peak and is not a roofline)."""
    sec = dom_ms * 1e-3
    out = {"kernel": dom, "mean_launch_ms": dom_ms, "traffic": traffic,
           "hbm_traffic_frac": (traffic / sec / (HBM_PEAK_GBS * 1e9)) if traffic else None,
           "plan_bytes_per_launch": alg_bytes, "plan_bytes_rate_GBs": alg_bytes / sec / 1e9}
    if lds and lds.get("valu_instr"):
        achieved = lds["valu_instr"] / sec / 1e9
        lds_rate = lds["bytes"] / sec / 1e9
        out.update({"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_GINSTR,
                    "unit": "G wave64 VALU instr/s", "frac": achieved / VALU_PEAK_GINSTR,
                    "valu_instr_per_launch": lds["valu_instr"],
                    "valu_lane_utilization": lds.get("lane_util"),
                    "peak_basis": "1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction",
                    "lds": {"bytes_per_launch": lds["bytes"], "achieved_GBs": lds_rate,
                            "peak_GBs": LDS_PEAK_GBS, "frac": lds_rate / LDS_PEAK_GBS,
                            "active_frac": lds.get("active"),
                            "bank_conflict_frac": lds.get("conflict"),
                            "peak_basis": "256 CUs x 128 B/clk x 2.4 GHz"}})
    elif traffic:
        achieved = traffic / sec / 1e9
        out.update({"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS})
    else:
        out.update({"bound": "valu", "achieved": None, "peak": VALU_PEAK_GINSTR,
                    "unit": "G wave64 VALU instr/s", "frac": None})
    return out


def tier_of(deg, k, caps):
    wave, lds, wide = caps
    t = np.full(len(deg), -1, dtype=np.int64)
    ok = deg >= k - 1
    t[ok & (deg <= wave)] = 0
    t[ok & (deg > wave) & (deg <= lds)] = 1
    t[ok & (deg > lds) & (deg <= wide)] = 2
    t[ok & (deg > wide)] = 3
    return t


TIER_KERNEL = {0: "clique_wave", 1: "clique_lds_block", 2: "clique_lds_block_wide",
               3: "clique_hbm_block"}


def cpu_baseline(d, dg, budget_s):
    """Times the oracle on a seeded uniform random sample of roots (k = 4 and
    5) and the engine on the same roots; checks the counts agree."""
    from oracle import oracle as O

    deg = d.degrees()
    rng = np.random.default_rng(SEED)
    order = rng.permutation(d.n).astype(np.uint32)
    # Calibrate the sample size on a small sample, then size it to ~budget_s.
    size = 2048
    t0 = time.perf_counter()
    O.clique_list(d.offsets, d.adj, 5, np.sort(order[:size]))
    O.clique_list(d.offsets, d.adj, 4, np.sort(order[:size]))
    t_cal = max(time.perf_counter() - t0, 1e-3)
    size = int(max(256, min(d.n, size * budget_s / t_cal)))
    roots = np.sort(order[:size])
    t0 = time.perf_counter()
    c4 = O.clique_list(d.offsets, d.adj, 4, roots)
    c5 = O.clique_list(d.offsets, d.adj, 5, roots)
    t_cpu = time.perf_counter() - t0
    edges = int(deg[roots].sum())
    g4 = dg.clique(4, roots=roots)
    g5 = dg.clique(5, roots=roots)
    # Engine on the same roots (tiers built by the calls above; timed again).
    gpu_ms = 0.0
    for k in (4, 5):
        best = None
        for _ in range(3):
            dg.clique(k, roots=roots)
            best = dg.last_kernel_ms() if best is None else min(best, dg.last_kernel_ms())
        gpu_ms += best
    exact = (g4 == c4) and (g5 == c5)
    return {
        "value": edges / t_cpu,
        "unit": "edges/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"{len(roots)} roots drawn uniformly at random (seed {SEED}; {edges} oriented "
                  f"edges, {edges / d.m:.2%} of |E+|), k=4 and k=5 counted one after the other "
                  f"(so edges/s = sample edges / time of both passes), oracle merge-intersection "
                  f"plan, OpenMP dynamic",
        "seconds": round(t_cpu, 3),
        "gpu_same_roots": {"ms": gpu_ms, "edges_per_s": edges / (gpu_ms * 1e-3),
                           "note": "k=4 and k=5 single-k passes over the same roots, "
                                   "event-timed, best of 3"},
        "exact_vs_engine": bool(exact),
        "counts": {"k4": int(c4), "k5": int(c5)},
    }


def extra_tc(with_cpu):
    """Triangle counting on seeded G(n=16384, M=131072), degree-oriented."""
    g = G.er_graph(16384, 131072, 20240601)
    d = G.orient(g)
    out = {"workload": "TC on Erdos-Renyi n=16384, M=131072, degree-oriented DAG",
           "oriented_edges": d.m}
    with G.DeviceGraph(d, oriented=True) as dg:
        plan = dg.clique_plan_bytes(3)
        out["algorithmic_bytes"] = plan
        for name, var in (("edge", G.TC_EDGE), ("vertex", G.TC_VERTEX)):
            c = dg.tc(var)  # first call builds the COO sources / root tiers
            first = dg.last_kernel_ms()
            singles = []
            for _ in range(5):
                assert dg.tc(var) == c
                singles.append(dg.last_kernel_ms())
            reps = 512
            means = []
            for _ in range(3):  # best of 3: the first replays run at a lower clock
                assert dg.tc_repeat(var, reps) == c
                means.append(dg.last_kernel_ms())
            mean = min(means)
            out[name] = {
                "count": int(c), "first_call_ms": first,
                "single_launch_ms": min(singles),
                "mean_ms_back_to_back": mean, "reps": reps,
                "repeat": "hipGraph of 8 chains of 4 counts" if var == G.TC_EDGE else "stream",
                "edges_per_s": d.m / (mean * 1e-3),
                "hbm_roofline_frac": plan / (mean * 1e-3) / (HBM_PEAK_GBS * 1e9),
            }
        # Up to 32 counts per launch (grid y = count index).
        means = []
        for _ in range(3):
            c3 = dg.tc_repeat_fused(512)
            assert c3 == out["edge"]["count"]
            means.append(dg.last_kernel_ms())
        mean = min(means)
        out["edge_fused"] = {"count": int(c3), "mean_ms_back_to_back": mean, "reps": 512,
                             "repeat": "32 counts per launch",
                             "edges_per_s": d.m / (mean * 1e-3),
                             "hbm_roofline_frac": plan / (mean * 1e-3) / (HBM_PEAK_GBS * 1e9)}
        # The k-clique bitmap path at k = 3 (rows built per root, popcounts).
        assert dg.clique(3) == out["edge"]["count"]
        ms = []
        for _ in range(5):
            dg.clique(3)
            ms.append(dg.last_kernel_ms())
        out["bitmap_k3"] = {"single_launch_ms": min(ms),
                            "hbm_roofline_frac": plan / (min(ms) * 1e-3) / (HBM_PEAK_GBS * 1e9)}
    if with_cpu:
        from oracle import oracle as O
        t0 = time.perf_counter()
        want = O.tc(d.offsets, d.adj)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": d.m / t, "unit": "edges/s", "cores": O.num_threads(),
                               "kind": "port", "count": int(want),
                               "exact_vs_engine": bool(want == out["edge"]["count"] ==
                                                       out["vertex"]["count"])}
    return out


def extra_sched(scale):
    """Learned vs plain degree scheduling: 5-clique on R-MAT scale 22, plus the
    3-/4-vertex motif counts of the same graph."""
    from cost_model import CostModel
    g = G.rmat_graph(scale, 16, 20240522)
    d = G.orient(g)
    out = {"workload": f"5-clique + motifs on R-MAT scale {scale}, edge factor 16",
           "n": d.n, "oriented_edges": d.m, "max_out_degree": int(d.degrees().max())}
    model = CostModel.load_default(5)
    with G.DeviceGraph(g, oriented=False) as sym, G.DeviceGraph(d, oriented=True) as dag:
        t0 = time.perf_counter()
        cost = model.predict(sym, dag)
        out["model_seconds"] = time.perf_counter() - t0
        plain = dag.clique(5)
        learned = dag.clique(5, policy=G.SCHED_LEARNED, cost=cost)
        split = dag.clique(5, policy=G.SCHED_DEGREE_SPLIT)
        t_plain, t_learned, t_split = [], [], []
        for _ in range(2):
            assert dag.clique(5) == plain
            t_plain.append(dag.last_kernel_ms())
            assert dag.clique(5, policy=G.SCHED_DEGREE_SPLIT) == split
            t_split.append(dag.last_kernel_ms())
            assert dag.clique(5, policy=G.SCHED_LEARNED, cost=cost) == learned
            t_learned.append(dag.last_kernel_ms())
        out["k5"] = {"count_plain": int(plain), "count_learned": int(learned),
                     "count_degree_split": int(split),
                     "identical": bool(plain == learned == split),
                     "plain_ms": min(t_plain), "learned_ms": min(t_learned),
                     "degree_split_ms": min(t_split),
                     "speedup": min(t_plain) / min(t_learned),
                     # The model's own share: the same splitting driven by the
                     # degree estimate d+^2 instead of the MLP.
                     "speedup_vs_degree_split": min(t_split) / min(t_learned),
                     "learned_breakdown_ms": dict(dag.kernel_breakdown())}
        tri = dag.clique(3)
        deg = g.degrees().astype(np.int64)
        motif, _ = sym.pattern("Motif4", G.FORM_PATTERN)
        out["motifs"] = {"triangles": int(tri),
                         "wedges_induced": int((deg * (deg - 1) // 2).sum() - 3 * tri),
                         "motif4_induced": dict(zip(["3-star", "4-path", "tailed-triangle",
                                                     "4-cycle", "diamond", "4-clique"],
                                                    [int(x) for x in motif])),
                         "motif4_ms": sym.last_kernel_ms(),
                         "motif4_breakdown_ms": dict(sym.kernel_breakdown())}
    return out


def extra_patterns(scale_both, scale_counting):
    """P1-P8 (G2MinerSolver and PatternSolver): both formulations on R-MAT
    scale_both (counts must agree), the counting formulation and its BLOCK_OPT
    prefilter on scale_counting.  The direct enumeration of P3 takes 18 s at
    scale 12 and the counting formulation of P2 68 s at scale 16, so the rows
    stay at scales 11 and 12."""
    out = {"workload": f"P1-P8 on R-MAT (edge factor 16, seed 20240522): both formulations at "
                       f"scale {scale_both}, counting formulation at scale {scale_counting}"}
    pids = ["P1", "P2", "P3", "P4", "P5", "P6", "P7", "P8"]
    g = G.rmat_graph(scale_both, 16, 20240522)
    rows = {}
    with G.DeviceGraph(g, oriented=False) as sym:
        for pid in pids:
            _, direct = sym.pattern(pid, G.FORM_G2MINER)
            t_direct = sym.last_kernel_ms()
            _, counted = sym.pattern(pid, G.FORM_PATTERN)
            t_count = sym.last_kernel_ms()
            rows[pid] = {"count": int(direct), "identical": int(direct) == int(counted),
                         "g2miner_ms": t_direct, "pattern_ms": t_count}
    out[f"scale{scale_both}"] = rows
    print("[bench] patterns small scale done", file=sys.stderr, flush=True)
    g = G.rmat_graph(scale_counting, 16, 20240522)
    rows = {}
    with G.DeviceGraph(g, oriented=False) as sym:
        for pid in pids:
            _, counted = sym.pattern(pid, G.FORM_PATTERN)
            t_count = sym.last_kernel_ms()
            _, pref = sym.pattern(pid, G.FORM_BLOCK_PREFILTER)
            rows[pid] = {"count": int(counted), "identical": int(counted) == int(pref),
                         "pattern_ms": t_count, "block_prefilter_ms": sym.last_kernel_ms()}
    out[f"scale{scale_counting}"] = rows
    print("[bench] patterns done", file=sys.stderr, flush=True)
    return out


def extra_match(scale, big_scale, with_cpu=False):
    """Labeled matching: the reference's 200 HPRD queries as one batch (exact
    against the reference's known answers), and the named queries on a
    labeled R-MAT graph."""
    import workloads as W
    out = {}
    z = np.load(os.path.join(ROOT, "tests", "golden", "ceci_hprd.npz"))
    e = z["data_edges"]
    g = G.csr_from_edges(int(z["data_n"]), e[:, 0], e[:, 1], G.CSR_CLEAN)
    qs, eo, lo = [], 0, 0
    for qn, qm in zip(z["q_n"], z["q_m"]):
        qn, qm = int(qn), int(qm)
        qs.append((qn, z["q_edges"][eo:eo + 2 * qm].reshape(-1, 2), z["q_labels"][lo:lo + qn]))
        eo, lo = eo + 2 * qm, lo + qn
    want = z["expected"].astype(np.uint64)
    res = {"workload": "200 reference queries (query_dense_16) on HPRD, one batch"}
    with G.DeviceGraph(g, oriented=False, labels=z["data_labels"]) as dg:
        for name, kw in (("plain", dict(autotune=1, policy=G.SCHED_DEGREE)),
                         ("learned", dict(autotune=1, policy=G.SCHED_LEARNED)),
                         ("graph_replay", dict(autotune=0, chunk=8, use_graph=1))):
            t0 = time.perf_counter()
            c = dg.match_batch(qs, **kw)
            first = (time.perf_counter() - t0) * 1e3  # planning, trials, roots, launch
            ms = []
            for _ in range(3):
                assert np.array_equal(dg.match_batch(qs, **kw), c)
                ms.append(dg.last_kernel_ms())
            res[name] = {"ms": min(ms), "first_call_ms": first,
                         "exact_vs_reference": bool(np.array_equal(c, want)),
                         "embeddings_per_s": float(c.sum()) / (min(ms) * 1e-3)}
        res["embeddings"] = int(want.sum())
        res["learned_speedup"] = res["plain"]["ms"] / res["learned"]["ms"]
    out["hprd_batch"] = res
    print("[bench] hprd batch done", file=sys.stderr, flush=True)
    g, labels = W.labeled_graph(scale)
    named = dict(W.named_queries())
    res = {"workload": f"named queries on labeled R-MAT scale {scale} (8 labels)"}
    with G.DeviceGraph(g, oriented=False, labels=labels, n_labels=W.N_LABELS) as dg:
        for qname in ("triangle", "4-cycle", "diamond", "three-star"):
            row = {}
            for pol, pname in ((G.SCHED_DEGREE, "plain"), (G.SCHED_LEARNED, "learned")):
                t0 = time.perf_counter()
                c = dg.match_batch([named[qname]], autotune=1, policy=pol)
                first = (time.perf_counter() - t0) * 1e3
                c2 = dg.match_batch([named[qname]], autotune=1, policy=pol)
                assert int(c2[0]) == int(c[0])
                row[pname] = {"count": int(c[0]), "ms": dg.last_kernel_ms(), "first_call_ms": first}
            row["identical"] = row["plain"]["count"] == row["learned"]["count"]
            res[qname] = row
            print(f"[bench] {qname} done", file=sys.stderr, flush=True)
        four = ("triangle", "4-cycle", "diamond", "three-star")
        row = {"queries": list(four)}
        for pol, pname in ((G.SCHED_DEGREE, "plain"), (G.SCHED_LEARNED, "learned")):
            t0 = time.perf_counter()
            c = dg.match_batch([named[q] for q in four], autotune=1, policy=pol)
            first = (time.perf_counter() - t0) * 1e3
            c2 = dg.match_batch([named[q] for q in four], autotune=1, policy=pol)
            assert np.array_equal(c, c2)
            assert [int(x) for x in c] == [res[q][pname]["count"] for q in four]
            row[pname] = {"ms": dg.last_kernel_ms(), "first_call_ms": first}
        row["learned_speedup"] = row["plain"]["ms"] / row["learned"]["ms"]
        res["one_batch"] = row
    out["named"] = res
    # The fifth named query, the 6-vertex pyramid, at scale 16: it has 3.1e9
    # embeddings there and grows ~10x per two scales.
    g, labels = W.labeled_graph(16)
    res = {"workload": "pyramid-6 (wheel W6) on labeled R-MAT scale 16 (8 labels)"}
    with G.DeviceGraph(g, oriented=False, labels=labels, n_labels=W.N_LABELS) as dg:
        for pol, pname in ((G.SCHED_DEGREE, "plain"), (G.SCHED_LEARNED, "learned")):
            t0 = time.perf_counter()
            c = dg.match_batch([named["pyramid-6"]], autotune=1, policy=pol)
            first = (time.perf_counter() - t0) * 1e3
            res[pname] = {"count": int(c[0]), "ms": dg.last_kernel_ms(), "first_call_ms": first}
        res["identical"] = res["plain"]["count"] == res["learned"]["count"]
    out["pyramid6_s16"] = res
    print("[bench] pyramid-6 done", file=sys.stderr, flush=True)
    # The same batch of four on the scale of the learned-scheduling workload.
    g, labels = W.labeled_graph(big_scale)
    res = {"workload": f"triangle, 4-cycle, diamond, three-star as one batch on labeled "
                       f"R-MAT scale {big_scale} (8 labels)"}
    with G.DeviceGraph(g, oriented=False, labels=labels, n_labels=W.N_LABELS) as dg:
        counts = {}
        # plain / learned: the autotuner picks each query's probe variant;
        # *_binary force per-lane binary search, *_merge the chunked merge.
        for pname, kw in (("plain", dict(policy=G.SCHED_DEGREE)),
                          ("learned", dict(policy=G.SCHED_LEARNED)),
                          ("plain_binary", dict(policy=G.SCHED_DEGREE,
                                                set_variant=G.MATCH_SETV_BINARY)),
                          ("plain_merge", dict(policy=G.SCHED_DEGREE,
                                               set_variant=G.MATCH_SETV_MERGE))):
            t0 = time.perf_counter()
            c = dg.match_batch([named[q] for q in four], autotune=1, **kw)
            first = (time.perf_counter() - t0) * 1e3
            counts[pname] = [int(x) for x in c]
            res[pname] = {"ms": dg.last_kernel_ms(), "first_call_ms": first}
            print(f"[bench] large named batch {pname} {res[pname]['ms']:.1f} ms",
                  file=sys.stderr, flush=True)
        res["counts"] = counts["plain"]
        res["identical"] = all(c == counts["plain"] for c in counts.values())
        res["learned_speedup"] = res["plain"]["ms"] / res["learned"]["ms"]
        res["merge_vs_binary_speedup"] = res["plain_binary"]["ms"] / res["plain_merge"]["ms"]
    out["named_batch_large"] = res
    print("[bench] large named batch done", file=sys.stderr, flush=True)
    # Batched matching workload: 256 seeded connected queries of 4-10
    # vertices (8 labels) as one batch; the autotuner picks each query's
    # chunk and probe variant, and short batches run in budgeted passes.
    # batch_256 runs it at workloads.BATCH_SCALE; batch_256_oracle_scale at
    # the scale the oracle enumerates in bench time, with the oracle timed
    # on the same batch (cpu_baseline) and compared query by query.
    qs = W.batch_queries()

    def run_batch(bs):
        g, labels = W.labeled_graph(bs)
        res = {"workload": f"{len(qs)} seeded connected queries (4-10 vertices, 8 labels) on "
                           f"labeled R-MAT scale {bs}, one batch",
               "sizes": [int(q[0]) for q in qs]}
        counts = {}
        with G.DeviceGraph(g, oriented=False, labels=labels, n_labels=W.N_LABELS) as dg:
            for pname, pol in (("learned", G.SCHED_LEARNED), ("plain", G.SCHED_DEGREE)):
                t0 = time.perf_counter()
                c = dg.match_batch(qs, autotune=1, policy=pol)
                first = (time.perf_counter() - t0) * 1e3
                ms = dg.last_kernel_ms()
                counts[pname] = c
                res[pname] = {"ms": ms, "first_call_ms": first,
                              "embeddings_per_s": float(c.sum()) / (ms * 1e-3)}
                print(f"[bench] batch of {len(qs)} at scale {bs} {pname} {ms:.1f} ms",
                      file=sys.stderr, flush=True)
        res["identical"] = bool(np.array_equal(counts["plain"], counts["learned"]))
        res["counts"] = [int(x) for x in counts["learned"]]
        res["embeddings"] = int(counts["learned"].sum())
        res["learned_speedup"] = res["plain"]["ms"] / res["learned"]["ms"]
        return res, g, labels, counts["learned"]

    out["batch_256"] = run_batch(W.BATCH_SCALE)[0]
    res, g, labels, got = run_batch(W.BATCH_ORACLE_SCALE)
    if with_cpu:
        # The oracle's CECI restatement over the same batch, parallel across
        # the queries (each query sequential, as in the reference).
        from oracle import oracle as O
        t0 = time.perf_counter()
        want = O.ceci_batch(g.offsets, g.adj, labels, qs)
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": float(want.sum()) / t, "unit": "embeddings/s",
                               "seconds": t, "cores": O.num_threads(), "kind": "port",
                               "parallel": "across queries",
                               "exact_vs_engine": bool(np.array_equal(want, got))}
        print(f"[bench] batch of {len(qs)} cpu oracle {t:.1f} s", file=sys.stderr, flush=True)
    out["batch_256_oracle_scale"] = res
    # Uneven batch: seeded random connected queries (random spanning tree plus
    # each other pair with probability 0.5), 3-6 vertices, 8 labels.
    rs = W.RANDOM_BATCH_SCALE
    qs = W.random_queries(32, W.SEEDS["batch"], min_size=3, max_size=6, extra_edge_prob=0.5)
    g, labels = W.labeled_graph(rs)
    res = {"workload": f"32 seeded random connected queries (3-6 vertices, 8 labels) on "
                       f"labeled R-MAT scale {rs}, one batch"}
    with G.DeviceGraph(g, oriented=False, labels=labels, n_labels=W.N_LABELS) as dg:
        counts = {}
        for pname, pol in (("plain", G.SCHED_DEGREE), ("learned", G.SCHED_LEARNED)):
            t0 = time.perf_counter()
            c = dg.match_batch(qs, autotune=1, policy=pol)
            first = (time.perf_counter() - t0) * 1e3
            c2 = dg.match_batch(qs, autotune=1, policy=pol)
            assert np.array_equal(c, c2)
            counts[pname] = c
            res[pname] = {"ms": dg.last_kernel_ms(), "first_call_ms": first}
        res["identical"] = bool(np.array_equal(counts["plain"], counts["learned"]))
        res["embeddings"] = int(counts["plain"].sum())
        res["learned_speedup"] = res["plain"]["ms"] / res["learned"]["ms"]
    out["random_batch"] = res
    print("[bench] random batch done", file=sys.stderr, flush=True)
    return out


def copy_calibration(mib):
    """A plain copy of known size: src read once, dst written once."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extras", default="tc,sched,match,patterns",
                    help="comma list of extra workloads: tc, sched, match, patterns (or none)")
    ap.add_argument("--sched-scale", type=int, default=22)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (to scale "
                         "profiler byte counters against a known traffic) and exit")
    ap.add_argument("--match-scale", type=int, default=18)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); splitting "
              "the roots over several GPUs (contiguous root ranges, one all-reduce of "
              "the uint64 counts) is not built.", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    g = G.rmat_graph(args.scale, 16, SEED)
    d = G.orient(g)
    t_prep = time.perf_counter() - t0
    deg = d.degrees()
    dg = G.DeviceGraph(d, oriented=True)
    ks = (4, 5)
    # One step = the 4- and 5-clique counts of every root, from one fused
    # pass (gpm_clique_all builds each root's rows once and counts k = 3, 4
    # and 5 from them).
    counts = {}
    for _ in range(max(args.warmup, 1)):
        c3, counts[4], counts[5] = dg.clique_all()

    # Timed region: every call is synchronous (ends with a stream sync).
    kern_ms = {}
    kern_launches = {}
    call_ms = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        got = dg.clique_all()
        assert got[1:] == (counts[4], counts[5]), "count changed between runs"
        call_ms += dg.last_kernel_ms()
        for name, ms in dg.kernel_breakdown():
            key = f"k45:{name}"
            kern_ms[key] = kern_ms.get(key, 0.0) + ms
            kern_launches[key] = kern_launches.get(key, 0) + 1
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / args.steps
    # One fused pass visits every oriented edge once as the first edge of its
    # cliques and yields both counts: value = |E+| per pass per second.
    value = d.m * args.steps / wall

    # The same counts as two separate passes (one per k), for reference.
    separate = {}
    for k in ks:
        assert dg.clique(k) == counts[k]
        best = None
        for _ in range(max(1, min(args.steps, 3))):
            dg.clique(k)
            best = dg.last_kernel_ms() if best is None else min(best, dg.last_kernel_ms())
        separate[f"k{k}_ms"] = best

    # Roofline of the dominant kernel: its roots' bytes under the reference's
    # k = 5 plan (which contains the k = 4 plan's set operations).
    dom = max(kern_ms, key=kern_ms.get)
    dom_k = 5
    dom_name = dom.split(":", 1)[1]
    caps = G.clique_tier_caps()
    tiers = tier_of(deg, dom_k, caps)
    tier_id = [t for t, n in TIER_KERNEL.items() if n == dom_name][0]
    work = dg.clique_root_work(dom_k)
    alg_bytes = 4 * int(work[tiers == tier_id].sum())
    dom_ms = kern_ms[dom] / kern_launches[dom]
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    plan_total = {k: dg.clique_plan_bytes(k) for k in ks}

    traffic = measured_traffic(dom, args.scale)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded R-MAT generated on the host)",
        "config": {
            "workload": f"k-clique k=4,5 on R-MAT scale {args.scale}, edge factor 16, "
                        "a,b,c,d=.57,.19,.19,.05, symmetrised+dedup, degree-oriented DAG; "
                        "both counts from one fused pass per step",
            "n": d.n, "oriented_edges": d.m, "max_out_degree": int(deg.max()),
            "seed": SEED, "parallelism": "single GPU",
        },
        "counts": {"k3": int(c3), "k4": int(counts[4]), "k5": int(counts[5])},
        # The reference's CFSolver runs one edge-rooted pass per k: the two
        # counts root at 2|E+| edges per step.
        "two_count_edges_per_s": 2 * d.m * args.steps / wall,
        "matches_per_s": (counts[4] + counts[5]) / (call_ms * 1e-3 / args.steps),
        "separate_passes_ms": separate,
        "kernel_ms_per_step": {k: v / args.steps for k, v in sorted(kern_ms.items())},
        "roofline": roofline(dom, dom_ms, alg_bytes, traffic, measured_lds(dom, args.scale)),
        "plan_bytes": {f"k{k}": int(v) for k, v in plan_total.items()},
        "prep_seconds": round(t_prep, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(d, dg, args.cpu_seconds)
    dg.close()
    extras = [e for e in args.extras.split(",") if e and e != "none"]
    out["extra"] = {}
    for e in extras:
        t0 = time.perf_counter()
        if e == "tc":
            out["extra"]["tc"] = extra_tc(not args.no_cpu)
        elif e == "sched":
            out["extra"]["sched"] = extra_sched(args.sched_scale)
        elif e == "match":
            out["extra"]["match"] = extra_match(args.match_scale, args.sched_scale,
                                                not args.no_cpu)
        elif e == "patterns":
            out["extra"]["patterns"] = extra_patterns(11, 12)
        print(f"[bench] extra {e} done in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
              flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
