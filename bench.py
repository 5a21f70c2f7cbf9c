#!/usr/bin/env python3
"""Benchmark: uint32 inclusive scan (chained scan with decoupled lookback) at
n = 2^28 full-range values on one MI355X, as GB/s at 8 algorithmic bytes per
element (4 read + 4 write) and as a fraction of the 8 TB/s HBM peak.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs all|none|a,b,..] [--no-cpu-baseline]

One step = one scan of the whole 2^28-element input (1 GiB in, 1 GiB out,
separate buffers, input resident in HBM). Prints ONE JSON line on stdout;
progress goes to stderr. The CPU baseline is the serial host scan (the
reference's validator loop restated in C) timed on this machine's host.

Extra legs (reported under "extras"): csdl_exclusive, rts, csdldf (stall
periods 0/7/8, one workgroup per tile and persistent), copy (the read-n + write-n ceiling), ceiling
(copies and one-way streams of the same 2 GiB with other access patterns, the
runtime's copy and fill), vendor (rocPRIM's device
scan, the reference's CUB comparator on ROCm), n2p20 (latency at 2^20),
segmented (4096 log-normal segments in one graph-launched pass) and
segmented_loop (the same as one launch per segment), segmented_fp and
segmented_fp_loop (the same segments over FP32: one segmented FP launch, and
one launch per segment), fp (n = 2^27 FP32/BF16 with and without compensated carries,
errors vs FP64), tuner (held-out regret of the learned cost model).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
BYTES_PER_ELEM = 8
ALL_LEGS = ["csdl_exclusive", "rts", "csdldf", "copy", "ceiling", "vendor", "n2p20", "segmented", "segmented_loop",
            "segmented_fp", "segmented_fp_loop", "fp",
            "struct", "occupancy", "survey", "tuner"]


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cpu() -> dict:
    """The host CPU as lscpu reports it, the logical CPUs the OS shows, and the
    CPUs this process may run on (its affinity mask)."""
    info = {"model": cpu_model(), "logical_cpus": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "logical_cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, ValueError):
        pass
    try:
        info["cpus_allowed"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["cpus_allowed"] = os.cpu_count()
    return info


def gbps(n: int, ms: float, bytes_per_elem: float = BYTES_PER_ELEM) -> float:
    return n * bytes_per_elem / (ms * 1e-3) / 1e9


def load_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary)."""
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return None


class Legs:
    """The extra measurements, each on the same resident 2^28 input."""

    def __init__(self, gps, torch, oracle, np, d_in, d_out, n):
        self.gps, self.torch, self.oracle, self.np = gps, torch, oracle, np
        self.d_in, self.d_out, self.n = d_in, d_out, n

    def _timed(self, handle, inclusive=True, iters=20):
        gps, n = self.gps, self.n
        for _ in range(3):
            handle.scan(self.d_in, self.d_out, n, inclusive)
        t = handle.time(self.d_in, self.d_out, n, inclusive, iters=iters)
        if gps.validate_scan_u32(self.d_in, self.d_out, n, inclusive) or handle.device_error():
            raise RuntimeError("validation failed")
        return statistics.median(t)

    def csdl_exclusive(self, out):
        with self.gps.ScanHandle(self.gps.ALGO_CSDL, self.n) as h:
            out["csdl_exclusive_gbps"] = round(gbps(self.n, self._timed(h, inclusive=False)), 1)

    def rts(self, out):
        with self.gps.ScanHandle(self.gps.ALGO_RTS, self.n) as h:
            m = self._timed(h)
            out["rts_inclusive_gbps"] = round(gbps(self.n, m), 1)
            out["rts_physical_gbps_12B"] = round(gbps(self.n, m, 12), 1)

    def csdldf(self, out):
        gps = self.gps
        schedules = (("csdldf", gps.SCHEDULE_TILE, {}),
                     ("csdldf_pp", gps.SCHEDULE_PERSISTENT, dict(block_threads=1024, vec_per_thread=8)))
        for name, schedule, shape in schedules:
            for period in (0, 7, 8):
                with gps.ScanHandle(gps.ALGO_CSDLDF, self.n, stall_period=period, schedule=schedule, **shape) as h:
                    for _ in range(3):
                        h.inclusive(self.d_in, self.d_out, self.n)
                    h.reset_counters()
                    m = self._timed(h, iters=20)
                    c = h.counters()
                    launches = 23  # 3 warm-up + 20 timed since the reset
                    out[f"{name}_stall{period}_gbps"] = round(gbps(self.n, m), 1)
                    out[f"{name}_stall{period}_per_launch"] = {k: round(v / launches, 1) for k, v in c.items()}
                    out[f"{name}_stall{period}_tiles"] = -(-self.n // h.tile)

    def copy(self, out):
        gps, torch = self.gps, self.torch
        st = torch.cuda.current_stream()
        for _ in range(3):
            gps.copy_u32(self.d_out, self.d_in, self.n)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(20):
            gps.copy_u32(self.d_out, self.d_in, self.n)
        e1.record(st)
        torch.cuda.synchronize()
        out["copy_ceiling_gbps"] = round(gbps(self.n, e0.elapsed_time(e1) / 20), 1)

    def ceiling(self, out):
        """The memory-ceiling probes on the scan's 2 GiB of traffic: copies of
        the 1 GiB input, and read-only / write-only streams of 2 GiB."""
        from gpuprefixsums_amd import ceiling
        torch = self.torch
        st = torch.cuda.current_stream()
        nbytes = self.n * 4
        big = torch.empty(2 * nbytes // 4, dtype=torch.int32, device="cuda")
        need = max(ceiling.scratch_bytes(k, nbytes) for k in range(ceiling.KINDS))
        scratch = torch.zeros(need // 4, dtype=torch.int32, device="cuda")
        res = {}
        for kind in range(ceiling.KINDS):
            if kind in ceiling.COPY_KINDS:
                src, dst, moved = self.d_in, self.d_out, nbytes
            elif kind == ceiling.READ_1024X1:
                src, dst, moved = big, self.d_out, 2 * nbytes
            else:
                src, dst, moved = None, big, 2 * nbytes
            for _ in range(3):
                ceiling.run(kind, src, dst, moved, scratch)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(20):
                ceiling.run(kind, src, dst, moved, scratch)
            e1.record(st)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 20
            traffic = 2 * nbytes  # a copy reads and writes nbytes; a stream moves 2 * nbytes one way
            entry = {"gbps": round(traffic / (ms * 1e-3) / 1e9, 1),
                     "frac": round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}
            if kind in ceiling.COPY_KINDS:
                entry["exact"] = bool(torch.equal(self.d_out, self.d_in))
                self.d_out.zero_()
            res[ceiling.name(kind)] = entry
        if int(scratch[0].item()) != 0:
            raise RuntimeError("ticket word not re-armed")
        del big, scratch
        out["ceiling_probes_2GiB"] = res

    def vendor(self, out):
        """rocPRIM's device scan (the reference's CUB comparator on ROCm) on the
        same input, separate buffers, inclusive, checked before it is reported."""
        gps, torch = self.gps, self.torch
        from gpuprefixsums_amd import vendor
        tb = vendor.vendor_scan_temp_bytes(self.n, True)
        temp = torch.empty(max(tb, 4), dtype=torch.uint8, device="cuda")
        for _ in range(3):
            vendor.vendor_scan_u32(self.d_in, self.d_out, self.n, temp, tb, True)
        ms = vendor.vendor_time_scan_u32(self.d_in, self.d_out, self.n, temp, tb, True, iters=20)
        torch.cuda.synchronize()
        errors = gps.validate_scan_u32(self.d_in, self.d_out, self.n, True)
        mean = sum(ms) / len(ms)
        out["rocprim_inclusive"] = {"gbps": round(gbps(self.n, mean), 1), "kernel_ms_mean": round(mean, 5),
                                    "validation_errors": errors, "temp_bytes": tb}

    def _c3_segments(self):
        """C3's 4096 segments: log-normal lengths (median 4096, sigma 1.5,
        clipped to [1, 2^22]), seed 3."""
        oracle, np, torch = self.oracle, self.np, self.torch
        S = 4096
        off = oracle.gen_lognormal_offsets(3, S, median=4096.0, sigma=1.5, lo=1, hi=1 << 22)
        return S, off, int(off[-1]), torch.from_numpy(off.view(np.int64)).to("cuda")

    def _events(self, fn, iters):
        torch = self.torch
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(iters):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / iters

    def segmented(self, out):
        """C3: 4096 log-normal segments of values uniform in [0, 1023] (seed 3),
        exclusive, in one launch: launched directly, and replayed from a
        hipGraph (the form a serving loop would use)."""
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        S, off, n, d_off = self._c3_segments()
        x = torch.from_numpy(oracle.gen_u32(3, n, 10).view(np.int32)).to("cuda")
        a, b = torch.empty_like(x), torch.empty_like(x)
        with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
            for _ in range(3):
                h.segmented(x, b, d_off, S, n, False)
            direct_ms = self._events(lambda: h.segmented(x, b, d_off, S, n, False), 20)
            with h.segmented_graph(x, a, d_off, S, n, False) as g:
                for _ in range(3):
                    g.launch()
                graph_ms = self._events(g.launch, 20)
            if not torch.equal(a, b) or h.device_error():
                raise RuntimeError("segmented outputs differ")
        nbytes = 8 * n + 8 * (S + 1)
        out["segmented"] = {"segments": S, "n_total": n, "graph_ms": round(graph_ms, 4),
                            "graph_gbps": round(nbytes / graph_ms / 1e6, 1), "direct_ms": round(direct_ms, 4),
                            "direct_gbps": round(nbytes / direct_ms / 1e6, 1), "bytes": nbytes}

    def segmented_loop(self, out):
        """C3's segments as one plain scan launch per segment (the baseline the
        batched kernel replaces), checked equal to the batched result."""
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        S, off, n, d_off = self._c3_segments()
        x = torch.from_numpy(oracle.gen_u32(3, n, 10).view(np.int32)).to("cuda")
        a, b = torch.empty_like(x), torch.empty_like(x)
        with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
            h.segmented(x, a, d_off, S, n, False)
            h.segmented_per_segment(x, b, off, False)
            torch.cuda.synchronize()
            loop_ms = self._events(lambda: h.segmented_per_segment(x, b, off, False), 3)
            if not torch.equal(a, b) or h.device_error():
                raise RuntimeError("segmented outputs differ")
        nbytes = 8 * n + 8 * (S + 1)
        r = {"per_segment_launches_ms": round(loop_ms, 3), "per_segment_gbps": round(nbytes / loop_ms / 1e6, 1)}
        if "segmented" in out:
            r["speedup_of_graph"] = round(loop_ms / out["segmented"]["graph_ms"], 1)
        out["segmented_loop"] = r

    def segmented_fp(self, out):
        """C3's segments over FP32 ~ N(0,1) values (seed 4), compensated
        inclusive, in one segmented FP launch; errors against the FP64 oracle."""
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        S, off, n, d_off = self._c3_segments()
        xh = oracle.gen_f32_normal(4, n)
        x = torch.from_numpy(xh).to("cuda")
        a = torch.empty_like(x)
        with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
            for _ in range(3):
                h.segmented_fp(x, a, d_off, S, n)
            batched_ms = self._events(lambda: h.segmented_fp(x, a, d_off, S, n), 20)
            if h.device_error():
                raise RuntimeError("device error in the segmented FP leg")
        y64 = oracle.segmented_scan_fp_f64(xh, off)
        s64 = oracle.segmented_scan_fp_f64(xh, off, absolute=True)
        e = np.abs(a.cpu().numpy().astype(np.float64) - y64)
        rel = e / np.maximum(np.abs(y64), 1.0)
        norm = e / np.maximum(s64, np.finfo(np.float64).tiny)
        nbytes = 8 * n + 8 * (S + 1)
        out["segmented_fp"] = {
            "segments": S, "n_total": n, "batched_ms": round(batched_ms, 4),
            "batched_gbps": round(nbytes / batched_ms / 1e6, 1),
            "max_rel_err": float(rel.max()), "median_rel_err": float(np.median(rel)),
            "max_normwise_err": float(norm.max()),
            "error_measures": "rel_i = |y_i - y64_i| / max(|y64_i|, 1); normwise_i = |y_i - y64_i| / S_i with "
                              "S_i = sum of |x_j| from the segment head through i"}

    def segmented_fp_loop(self, out):
        """C3's FP32 segments as one compensated FP scan launch per segment,
        with the same error measures."""
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        S, off, n, d_off = self._c3_segments()
        xh = oracle.gen_f32_normal(4, n)
        x = torch.from_numpy(xh).to("cuda")
        b = torch.empty_like(x)
        with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
            h.fp_per_segment(x, b, off)
            torch.cuda.synchronize()
            loop_ms = self._events(lambda: h.fp_per_segment(x, b, off), 3)
            if h.device_error():
                raise RuntimeError("device error in the per-segment FP leg")
        y64 = oracle.segmented_scan_fp_f64(xh, off)
        s64 = oracle.segmented_scan_fp_f64(xh, off, absolute=True)
        e = np.abs(b.cpu().numpy().astype(np.float64) - y64)
        nbytes = 8 * n + 8 * (S + 1)
        r = {"per_segment_launches_ms": round(loop_ms, 3), "per_segment_gbps": round(nbytes / loop_ms / 1e6, 1),
             "max_rel_err": float((e / np.maximum(np.abs(y64), 1.0)).max()),
             "max_normwise_err": float((e / np.maximum(s64, np.finfo(np.float64).tiny)).max())}
        if "segmented_fp" in out:
            r["speedup_of_batched"] = round(loop_ms / out["segmented_fp"]["batched_ms"], 1)
        out["segmented_fp_loop"] = r

    def fp(self, out):
        """n = 2^27 inclusive FP scans: FP32 ~ N(0,1) (seed 4) and BF16 with
        magnitudes 10^U(-4,4) and random sign (seed 5); carries as two-float
        TwoSum pairs ("compensated") or as plain FP32 ("plain"); GB/s (8 B/elem
        FP32, 6 B/elem BF16 in + FP32 out) and two error measures against the
        FP64 oracle: max / median rel_i = |y_i - y64_i| / max(|y64_i|, 1), and
        max normwise_i = |y_i - y64_i| / S_i with S_i = sum_{j<=i} |x_j|."""
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        n = 1 << 27
        st = torch.cuda.current_stream()
        res = {}
        dout = torch.empty(n, dtype=torch.float32, device="cuda")
        with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
            for kind in ("f32", "bf16"):
                if kind == "f32":
                    x = oracle.gen_f32_normal(4, n)
                    y64, s64 = oracle.scan_f32_f64(x), oracle.abs_scan_f32_f64(x)
                    din = torch.from_numpy(x).cuda()
                    fn, bpe = h.inclusive_f32, 8
                else:
                    x = oracle.gen_bf16_logmag(5, n)
                    y64, s64 = oracle.scan_bf16_f64(x), oracle.abs_scan_bf16_f64(x)
                    din = torch.from_numpy(x.view(np.int16)).cuda()
                    fn, bpe = h.inclusive_bf16_f32, 6
                den = np.maximum(np.abs(y64), 1.0)
                sden = np.maximum(s64, np.finfo(np.float64).tiny)
                for comp in (gps.FP_PAIR, gps.FP_PLAIN):
                    for _ in range(3):
                        fn(din, dout, n, comp)
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(st)
                    for _ in range(10):
                        fn(din, dout, n, comp)
                    e1.record(st)
                    torch.cuda.synchronize()
                    ms = e0.elapsed_time(e1) / 10
                    err = np.abs(dout.cpu().numpy().astype(np.float64) - y64)
                    rel = err / den
                    mode = {gps.FP_PAIR: "compensated", gps.FP_PLAIN: "plain"}[comp]
                    res[f"{kind}_{mode}"] = {
                        "gbps": round(n * bpe / ms / 1e6, 1), "max_rel_err": float(rel.max()),
                        "median_rel_err": float(np.median(rel)), "max_normwise_err": float((err / sden).max())}
                    del err, rel
                del x, y64, s64, din, den, sden
            if h.device_error():
                raise RuntimeError("device error in fp leg")
        out["fp_n2p27"] = res

    def struct(self, out):
        """Member-wise scan of n/4 = 2^26 four-member uint32 structs over the
        same 1 GiB input (inclusive and exclusive), checked member by member
        against the oracle; 32 B per struct."""
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        ns = self.n // 4
        st = torch.cuda.current_stream()
        res = {}
        with gps.ScanHandle(gps.ALGO_CSDL, ns) as h:
            for inclusive in (True, False):
                for _ in range(3):
                    h.scan_u32x4(self.d_in, self.d_out, ns, inclusive)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                for _ in range(20):
                    h.scan_u32x4(self.d_in, self.d_out, ns, inclusive)
                e1.record(st)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / 20
                want = oracle.scan_u32x4(self.d_in.cpu().numpy().view(np.uint32), inclusive)
                got = self.d_out.cpu().numpy().view(np.uint32).reshape(-1, 4)
                if not np.array_equal(got, want) or h.device_error():
                    raise RuntimeError("struct scan differs from the oracle")
                res["inclusive" if inclusive else "exclusive"] = {"gbps": round(gbps(ns, ms, 32), 1),
                                                                  "kernel_ms": round(ms, 5)}
        out["struct_u32x4_n2p26"] = res

    def occupancy(self, out):
        """Occupancy estimate (one workgroup launched per tile; those that took a
        tile are counted) for the persistent u32 scan at 2^28 and the struct
        scan at 2^26 structs, against the resident grid the normal launch uses;
        the oversubscribed scans are checked like the others."""
        gps, np = self.gps, self.np
        res = {}
        with gps.ScanHandle(gps.ALGO_CSDL, self.n) as h:
            m, r = h.estimate_occupancy(self.d_in, self.d_out, self.n)
            if gps.validate_scan_u32(self.d_in, self.d_out, self.n, True) or h.device_error():
                raise RuntimeError("oversubscribed scan failed validation")
            res["u32_n2p28"] = {"workgroups_that_took_a_tile": m, "resident_grid": r, "tiles": -(-self.n // h.tile)}
            m, r = h.estimate_occupancy(self.d_in, self.d_out, self.n // 4, struct_scan=True)
            if h.device_error():
                raise RuntimeError("device error in the struct occupancy run")
            res["struct_n2p26"] = {"workgroups_that_took_a_tile": m, "resident_grid": r,
                                   "tiles": -(-(self.n // 4) // 8192)}
        out["occupancy"] = res

    def survey(self, out):
        """Survey networks (serial, wave / block Kogge-Stone, Sklansky,
        Brent-Kung, raking, DPP, the register tile) scanning independent
        segments of their own length over 2^26 of the input, inclusive: GB/s
        at 8 B per element, each result checked against the oracle."""
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        n = 1 << 26
        x = self.d_in[:n]
        y = self.d_out[:n]
        xh = x.cpu().numpy().view(np.uint32)
        refs = {}
        st = torch.cuda.current_stream()
        res = {}
        for name, alg in gps.SURVEY_NETWORKS.items():
            seg = gps.survey_segment(alg)
            for _ in range(2):
                gps.survey_scan_u32(alg, x, y, n, True)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(10):
                gps.survey_scan_u32(alg, x, y, n, True)
            e1.record(st)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 10
            if seg not in refs:
                refs[seg] = oracle.segmented_scan_u32(xh, np.arange(0, n + 1, seg, dtype=np.uint64), True)
            if not np.array_equal(y.cpu().numpy().view(np.uint32), refs[seg]):
                raise RuntimeError(f"survey network {name} differs from the oracle")
            res[name] = {"segment": seg, "gbps": round(gbps(n, ms), 1), "kernel_ms": round(ms, 4)}
        out["survey_n2p26"] = res

    def tuner(self, out):
        """Learned tuner: regret of the tuner's choices (MLP cost model, inference
        on the GPU with the library's kernels, gated against the default) on the
        held-out workloads of the 512-workload sweep and of the decoupled-
        fallback (stall) sweep, against the exhaustively measured best; and,
        without the model, the regret of the library's per-size default, of its
        fixed 1024x8 default and of the best static rule on the same workloads.
        Then 16 held-out workloads (n <= 2^26, <= 2048 segments) are re-timed
        here, every candidate of each: the drift of those timings against the
        committed ones, and the regrets on the fresh timings."""
        from gpuprefixsums_amd.tuner import DATA, DEFAULT_MODEL
        from gpuprefixsums_amd.tuner.evaluate import best_static, load_sweeps, regrets, summarize
        from gpuprefixsums_amd.tuner.harness import Runner
        from gpuprefixsums_amd.tuner.model import CostModel
        times_path, stall_path = DATA / "c5_times_mi355x.jsonl", DATA / "stall_times_mi355x.jsonl"
        if not times_path.exists() or not DEFAULT_MODEL.exists():
            out["tuner"] = None
            return
        times, train, held, s_train, s_held = load_sweeps(times_path, stall_path)
        static = best_static(train + s_train, times)
        model = CostModel.load(DEFAULT_MODEL)
        r_mlp, r_def, r_stat, rows = regrets(model, held, times, static)
        res = {"held_out": len(held), "regret_mlp": r_mlp, "regret_default": r_def,
               "regret_fixed_default": summarize([r["fixed_regret"] for r in rows]),
               "regret_best_static": r_stat, "best_static": static,
               "source": "tuner/data/c5_times_mi355x.jsonl (exhaustive harness timings on MI355X)"}
        if s_held:
            s_mlp, s_def, s_stat, _ = regrets(model, s_held, times, static)
            res["stall_sweep"] = {"held_out": len(s_held), "regret_mlp": s_mlp, "regret_default": s_def,
                                  "regret_best_static": s_stat}
        subset = [w for w in held if w.n <= 1 << 26 and w.segments <= 2048][:16]
        if subset:
            runner = Runner(max_n=max(w.n for w in subset), max_segments=max(w.segments for w in subset))
            fresh, ratios = {}, []
            for w in subset:
                fresh[w.wid] = runner.time_all(w)
                ratios += [fresh[w.wid][c] / times[w.wid][c] for c in fresh[w.wid] if c in times[w.wid]]
            f_mlp, f_def, _, _ = regrets(model, subset, fresh, static)
            np = self.np
            dev = np.abs(np.array(ratios) - 1.0)
            res["retimed"] = {"workloads": [w.wid for w in subset], "timings": len(ratios),
                              "drift_median": float(np.median(dev)), "drift_max": float(dev.max()),
                              "regret_mlp_fresh": f_mlp, "regret_default_fresh": f_def}
            del runner
            self.torch.cuda.empty_cache()
        out["tuner"] = res

    def n2p20(self, out):
        gps, torch, oracle, np = self.gps, self.torch, self.oracle, self.np
        n1 = 1 << 20
        x1 = torch.from_numpy(oracle.gen_u32(1, n1, 8).view(np.int32)).to("cuda")
        o1 = torch.empty_like(x1)
        with gps.ScanHandle(gps.ALGO_CSDL, n1) as h1:
            for _ in range(5):
                h1.inclusive(x1, o1, n1)
            t = h1.time(x1, o1, n1, True, iters=50)
            out["n2p20_csdl_inclusive_us"] = round(statistics.median(t) * 1e3, 2)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--legs", default="all", help="all, none, or a comma list of " + ",".join(ALL_LEGS))
    ap.add_argument("--cpu-baseline", action=argparse.BooleanOptionalAction, default=True)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this library is single-GPU (the multi-GPU "
              "sharding of the scan is a design note, not built); run with --gpus 1.", file=sys.stderr)
        return 2
    legs = ALL_LEGS if args.legs == "all" else [] if args.legs == "none" else args.legs.split(",")
    unknown = [x for x in legs if x not in ALL_LEGS]
    if unknown:
        print(f"unknown legs {unknown}", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import gpuprefixsums_amd as gps
    import oracle

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    n = 1 << args.log2n
    K, W = args.steps, args.warmup
    seed = 2

    log(f"generating n=2^{args.log2n} full-range uint32 input (splitmix64 seed {seed}) on the host")
    x = oracle.gen_u32(seed, n, 32)
    d_in = torch.from_numpy(x.view(np.int32)).to("cuda")
    d_out = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    h = gps.ScanHandle(gps.ALGO_CSDL, n)
    for _ in range(W):
        h.inclusive(d_in, d_out, n)
    torch.cuda.synchronize()

    # Timed region: K scans back to back on torch's current stream (the
    # library launches on it for tensor arguments), bracketed by one pair of
    # hipEvents recorded on that stream, and by synchronisations for the
    # wall-clock figure. The average launch duration is the events' elapsed
    # time over K.
    st = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(st)
    for _ in range(K):
        h.inclusive(d_in, d_out, n)
    ev1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mean_ms = ev0.elapsed_time(ev1) / K
    errors = gps.validate_scan_u32(d_in, d_out, n, True)
    dev_err = h.device_error()
    if errors or dev_err:
        log(f"FAILED: {errors} validation errors, device error bits {dev_err}")
        return 1
    # The per-launch distribution, outside the timed region: K more scans,
    # each bracketed by hipEvents (the extra event between launches costs a
    # little, so these run slightly longer than the back-to-back average).
    ms = h.time(d_in, d_out, n, True, iters=K)
    med_ms = statistics.median(ms)
    value = K * n * BYTES_PER_ELEM / wall / 1e9
    log(f"csdl inclusive: {value:.1f} GB/s whole-job; kernel average {mean_ms * 1e3:.1f} us back to back "
        f"-> {gbps(n, mean_ms):.1f} GB/s; per-launch median {med_ms * 1e3:.1f} us")

    traffic = load_traffic(f"csdl_u32_n2p{args.log2n}")
    schedule = {1: "tile", 2: "persistent"}.get(h.schedule, str(h.schedule))
    roofline = {
        "bound": "hbm",
        "achieved": round(gbps(n, mean_ms), 1),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(gbps(n, mean_ms) / HBM_PEAK_GBPS, 4),
        "traffic": traffic.get("bytes_per_launch") if traffic else None,
        "kernel": f"csdl_pp_u32<{h.block_threads},{h.vec_per_thread},incl> ({schedule} schedule)",
        "algorithmic_bytes": n * BYTES_PER_ELEM,
        "kernel_ms_mean": round(mean_ms, 5),
        "timing": "kernel_ms_mean: hipEvent pair around the K timed launches, divided by K; the per-launch "
                  "figures below come from K further launches each bracketed by hipEvents",
        "per_launch_ms_mean": round(statistics.fmean(ms), 5),
        "per_launch_ms_median": round(med_ms, 5),
        "per_launch_ms_p5_p95": [round(float(np.percentile(ms, 5)), 5), round(float(np.percentile(ms, 95)), 5)],
    }

    extras = {}
    runner = Legs(gps, torch, oracle, np, d_in, d_out, n)
    for leg in legs:
        t = time.perf_counter()
        getattr(runner, leg)(extras)
        log(f"leg {leg} done in {time.perf_counter() - t:.1f} s")
    if "copy_ceiling_gbps" in extras:
        extras["csdl_frac_of_copy_ceiling"] = round(gbps(n, mean_ms) / extras["copy_ceiling_gbps"], 4)

    cpu = None
    if args.cpu_baseline:
        log("timing the serial host scan (oracle port of the reference's validator loop)")
        out = np.empty_like(x)
        runs = []
        for _ in range(3):
            t = time.perf_counter()
            oracle.lib().orc_scan_u32(x, out, n, 1)
            runs.append(time.perf_counter() - t)
        s = statistics.median(runs)
        host = host_cpu()
        cpu = {"value": round(n * BYTES_PER_ELEM / s / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"full workload n=2^{args.log2n} inclusive, serial scan on one core, median of 3 runs "
                         f"({s * 1e3:.0f} ms each) on {host['model']}",
               "host": host}
        # The blocked two-pass scan on every CPU this process is granted: the
        # affinity mask, further capped by OMP_NUM_THREADS where the launcher
        # sets the per-GPU CPU share there.
        threads = host["cpus_allowed"] or 1
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
        runs = []
        for _ in range(3):
            t = time.perf_counter()
            oracle.lib().orc_scan_u32_mt(x, out, n, 1, threads)
            runs.append(time.perf_counter() - t)
        s = statistics.median(runs)
        extras["cpu_parallel"] = {"value": round(n * BYTES_PER_ELEM / s / 1e9, 3), "unit": "GB/s",
                                  "threads": threads, "kind": "port",
                                  "sample": "blocked two-pass scan (tile sums, serial spine, rescan) on the CPUs "
                                            "granted to this process (affinity mask, capped by OMP_NUM_THREADS)"}
        log(f"cpu baseline: {json.dumps(cpu)}")

    line = {
        "metric": "scan GB/s & %HBM peak, u32 n=2^28; segmented batch vs plain; tuner regret",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"uint32 inclusive scan, CSDL, n=2^{args.log2n}, values uniform in "
                               f"[0, 2^32-1] (splitmix64 seed {seed}), separate 1 GiB in/out",
                   "n": n, "algo": "csdl", "tile": h.tile, "schedule": schedule,
                   "bytes_per_elem": BYTES_PER_ELEM},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    h.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
