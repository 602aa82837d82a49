"""Benchmark: read pairs/s mapping 2x150 paired-end reads to a 100 Mbp reference.

Workload: 100 Mbp in 20 contigs of 5 Mbp (seed 7) with ~30% copied 1-10 kbp
segments at 2% divergence; 1,000,000 pairs 2x150 (fragment Normal(300,30),
0.5% substitutions, 0.05% indels) generated by strobealign_amd/synth.py. One
step maps all pairs through the HIP library (seeding, lookups, chaining, NAMs,
extension on the GPU; decision replay and SAM formatting on the host) with the
reads parsed once and staged in HBM before timing and the index resident in
HBM. The SAM text is produced per chunk in host memory and then dropped.

Then the serving loop (sa_server_*) runs on the same resident index: requests
of log-uniform [1, 50000] pairs cut from the same reads arrive Poisson at 200/s
(400 requests), each timed from its scheduled arrival to its last SAM byte
(p50/p99; four client threads submit the requests as they arrive);
the same requests submitted at once give the saturated throughput of the request
loop. A second serving run follows the serving configuration's request shape:
one read length per request, uniform in [100, 300] bp (25 bp steps of
pre-generated 2xL read sets on the same reference), on an index built for read
length 250; its p50/p99 are the line's "latency_ms".

Prints one JSON line (see the contract in the repository docs) with a roofline
object for the dominant kernel (HIP-event timing on the engine's stream) and a
CPU baseline: the oracle (a C++ restatement of the reference) on the host's
cores over a bounded sample of the same reads.
"""
import argparse
import json
import math
import os
import queue
import random
import subprocess
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(ref, r1, r2, pairs, threads, workdir, out=os.devnull, t1_chunks=3, read_length=0):
    """Times the oracle on the first `pairs` records (pairs, or reads when r2 is
    None; index build excluded); its SAM goes to `out`. Returns (records/s,
    oracle timings, sample files)."""
    oracle = os.path.join(ROOT, "oracle", "build", "sa_oracle")
    if not os.path.exists(oracle):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    s1 = os.path.join(workdir, "cpu_1.fq")
    s2 = os.path.join(workdir, "cpu_2.fq") if r2 else None
    for src, dst in ((r1, s1), (r2, s2)):
        if not src:
            continue
        with open(src, "rb") as f, open(dst, "wb") as g:
            for i, line in enumerate(f):
                if i >= 4 * pairs:
                    break
                g.write(line)
    cmd = [oracle, "-t", str(threads), "--no-PG", "--time", "--t1-chunks", str(t1_chunks), "-o", out]
    if read_length:
        cmd += ["-r", str(read_length)]
    p = subprocess.run(cmd + [ref, s1] + ([s2] if s2 else []), check=True, capture_output=True, text=True)
    t = json.loads(p.stderr.strip().splitlines()[-1])
    return pairs / t["map_s"], t, (s1, s2)


def request_plan(n_pairs, n_req, rate, seed):
    """Poisson arrivals (rate per second) of requests of log-uniform [1, 50000]
    pairs, cut consecutively from the read set (wrapping): (arrival_s, first, count)."""
    rnd = random.Random(seed)
    plan, pos, t = [], 0, 0.0
    for _ in range(n_req):
        t += rnd.expovariate(rate)
        n = min(50_000, max(1, int(math.exp(rnd.uniform(0.0, math.log(50_000))))), n_pairs)
        if pos + n > n_pairs:
            pos = 0
        plan.append((t, pos, n))
        pos += n
    return plan


def serve(sa, ix, reads, threads, slots, plan, paced):
    """Runs a request plan through the serving loop. paced: each request arrives
    at its scheduled Poisson time and is handed to one of four submitting client
    threads (so a large request being packed does not hold back the arrivals
    behind it); otherwise all are submitted at once by four clients (saturation).
    Returns latencies in ms measured from each request's scheduled arrival to its
    last output byte (paced runs), the wall time from the first arrival to the last
    completion, server stats, and the pacer's and the submitters' largest delays
    behind the schedule (seconds)."""
    # the plain per-batch stages: the server's batches are short and of varying
    # shape (8 Mbases by default), which the fused pass, recorded and sized for a
    # batch shape, serves worse (measured: 0.93M against 3.2M pairs/s saturated)
    srv = sa.Server(ix, sa.options(threads=threads, slots=slots))
    lat = [None] * len(plan)
    q = queue.Queue()

    def waiter():
        while True:
            item = q.get()
            if item is None:
                return
            k, t, before_ms = item
            # library latency counts from the start of the submit call
            lat[k] = before_ms + srv.wait(t, keep_output=False)[1]
    ws = [threading.Thread(target=waiter) for _ in range(8)]
    for w in ws:
        w.start()
    t0 = time.perf_counter()
    lag = 0.0
    submit_delay = [0.0]

    def submit(k, t_arr):
        t_sub = time.perf_counter()
        ticket = srv.submit(plan[k][3] if len(plan[k]) > 3 else reads, plan[k][1], plan[k][2])
        submit_delay[0] = max(submit_delay[0], t_sub - t_arr)
        q.put((k, ticket, 1000.0 * (t_sub - t_arr)))
    if paced:
        sq = queue.Queue()

        def client():
            while True:
                item = sq.get()
                if item is None:
                    return
                submit(*item)
        cs = [threading.Thread(target=client) for _ in range(4)]
        for c in cs:
            c.start()
        for k, (ta, *_) in enumerate(plan):
            d = t0 + ta - time.perf_counter()
            if d > 0:
                time.sleep(d)
            lag = max(lag, -d)
            sq.put((k, t0 + ta))
        for _ in cs:
            sq.put(None)
        for c in cs:
            c.join()
    else:
        # a backlog from several clients at once: four submitting threads
        def client(ks):
            for k in ks:
                submit(k, t0)
        cs = [threading.Thread(target=client, args=(list(range(c, len(plan), 4)),)) for c in range(4)]
        for c in cs:
            c.start()
        for c in cs:
            c.join()
    for _ in ws:
        q.put(None)
    for w in ws:
        w.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return lat, wall, st, (lag, submit_delay[0])


def serve_varied(a, sa, synth, ctx, threads, slots):
    """Requests as the serving configuration describes them: Poisson arrivals,
    log-uniform [1, 50000] pairs, one read length per request drawn uniformly
    from [100, 300] (rounded to the 25 bp steps of the pre-generated read sets),
    on a 100 Mbp index built for read length 250. Returns the serving object."""
    lengths = list(range(100, 301, 25))
    w = synth.make_lengths(os.path.join(a.data_dir, "lengths"), lengths, 50_000)
    ix = sa.Index(ctx, w["ref"], read_length=250, device_build=True)
    rs = {L: sa.Reads(*w["sets"][L]) for L in lengths}
    for r in rs.values():
        r.upload(ix)
    rnd = random.Random(13)
    plan, t, pos = [], 0.0, {L: 0 for L in lengths}
    for _ in range(a.serve_varied):
        t += rnd.expovariate(a.serve_rate)
        n = min(50_000, max(1, int(math.exp(rnd.uniform(0.0, math.log(50_000))))))
        L = lengths[min(len(lengths) - 1, int(round((rnd.uniform(100, 300) - 100) / 25)))]
        if pos[L] + n > 50_000:
            pos[L] = 0
        plan.append((t, pos[L], n, rs[L]))
        pos[L] += n
    lat, wall, st, (lag, sub_delay) = serve(sa, ix, None, threads, slots, plan, True)
    _, sat_wall, _, _ = serve(sa, ix, None, threads, slots, plan, False)
    pairs = sum(p[2] for p in plan)
    out = {"requests": len(plan), "arrivals": "Poisson %.0f/s" % a.serve_rate,
           "read_lengths": "uniform [100, 300] per request, in 25 bp steps", "index_read_length": 250,
           "pairs_per_request": "log-uniform [1, 50000] (mean %.0f)" % (pairs / len(plan)),
           "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2), "max_ms": round(max(lat), 2),
           "served_pairs_per_s": round(pairs / wall, 1), "saturated_pairs_per_s": round(pairs / sat_wall, 1),
           "batches": st["batches"], "latency_from": "scheduled Poisson arrival to the last output byte", "pass": "plain",
           "driver_max_lag_ms": round(1000 * lag, 2), "max_submit_delay_ms": round(1000 * sub_delay, 2)}
    for r in rs.values():
        r.close()
    ix.close()
    print("serving (100-300 bp): p50 %.1f ms, p99 %.1f ms, saturated %.0f pairs/s" %
          (out["p50_ms"], out["p99_ms"], out["saturated_pairs_per_s"]), file=sys.stderr, flush=True)
    return out


# Counter profiles -> per-launch traffic and VALU figures (bench.py --summarize-pmc).
# Engine stages (the names the engine times) and the HIP kernels each launches
# once per batch: a stage's traffic per launch is the sum over its kernels.
# the kernels of each stage (helpers shared between stages, such as the counter
# clear k_zero_u32, are left out: their dispatch counts are not the stage's)
STAGE_KERNELS = {"k_hits": ("k_hits", "k_hits_grp"), "k_chain": ("k_chain_t1", "k_chain_tn", "k_chain"),
                 "k_seed": ("k_seed", "k_seed_wave"), "k_finalize": ("k_finalize",)}
COPY_BYTES = 1 << 30  # k_hbm_copy reads and writes 1 GiB per dispatch
VALU_INT_PEAK_TOPS = 39.3  # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, int32 ops
VALU_WAVE_INSTR_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions/s: 256 CUs x 4 SIMDs, 4 cycles each


def _short(kernel_name):
    return kernel_name.split("(")[0].replace("void ", "").replace("sa::", "").strip()


def pmc_stages(kernels):
    """Per-stage traffic per launch from the per-kernel counter entries: a stage
    launches once per batch, as often as its most frequent kernel."""
    stages = {}
    for stage, ks in STAGE_KERNELS.items():
        members = [k for k in kernels if k.split("<")[0] in ks]
        if not members or any("hbm_bytes" not in kernels[k] for k in members):
            continue
        launches = max(kernels[k]["dispatches"] for k in members)
        tot = lambda f: sum(kernels[k][f] * kernels[k]["dispatches"] for k in members) / launches
        st = {"launches": launches, "kernels": sorted(members), "hbm_bytes_per_launch": int(tot("hbm_bytes")),
              "dram_read_bytes": int(tot("dram_read_bytes")), "dram_write_bytes": int(tot("dram_write_bytes"))}
        if all("fetch_size_bytes_scaled" in kernels[k] and "write_size_bytes_scaled" in kernels[k] for k in members):
            st["fetch_write_size_bytes_scaled"] = int(tot("fetch_size_bytes_scaled") + tot("write_size_bytes_scaled"))
        stages[stage] = st
    return stages


def summarize_pmc(dbs, workload, batch_bases):
    """Per-kernel counter averages per dispatch from rocprofv3 --pmc databases
    (one pass per database), the stage traffic per launch, and the text table."""
    import sqlite3
    per = {}  # kernel -> counter -> {dispatch: value}
    dur = {}  # kernel -> {dispatch: ns}
    for db in dbs:
        c = sqlite3.connect(db)
        for kn, did, cn, val, st, en in c.execute(
                "select kernel_name, dispatch_id, counter_name, value, start, end from counters_collection"):
            k = _short(kn)
            d = per.setdefault(k, {}).setdefault(cn, {})
            d[did] = d.get(did, 0.0) + float(val)
            dur.setdefault(k, {})[(db, did)] = en - st
        c.close()

    def avg(k, cn):
        # a counter's per-dispatch mean, collected as the plain counter (summed
        # over its dimensions here) or as its _sum form
        d = per.get(k, {}).get(cn) or per.get(k, {}).get(cn[:-4] if cn.endswith("_sum") else cn + "_sum")
        return sum(d.values()) / len(d) if d else None

    def ndisp(k):
        return max((len(v) for v in per.get(k, {}).values()), default=0)

    def rd_bytes(k):  # every TCC/EA read request by size
        a, b, c_ = avg(k, "TCC_EA0_RDREQ_32B_sum"), avg(k, "TCC_EA0_RDREQ_64B_sum"), avg(k, "TCC_EA0_RDREQ_128B_sum")
        return None if a is None else 32 * a + 64 * b + 128 * c_

    def dram_rd(k):  # read requests destined for DRAM, in 32-byte units
        v = avg(k, "TCC_EA0_RDREQ_DRAM_32B_sum")
        return None if v is None else 32 * v

    def dram_wr(k):
        v = avg(k, "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum")
        return None if v is None else 32 * v

    def wr_bytes(k):
        a, b = avg(k, "TCC_EA0_WRREQ_sum"), avg(k, "TCC_EA0_WRREQ_64B_sum")
        return None if a is None else 32 * (a - b) + 64 * b

    cp = "k_hbm_copy"
    f_rd = COPY_BYTES / dram_rd(cp) if dram_rd(cp) else None
    f_wr = COPY_BYTES / dram_wr(cp) if dram_wr(cp) else None
    f_fetch = COPY_BYTES / 1024 / avg(cp, "FETCH_SIZE") if avg(cp, "FETCH_SIZE") else None
    f_write = COPY_BYTES / 1024 / avg(cp, "WRITE_SIZE") if avg(cp, "WRITE_SIZE") else None
    out = {"workload": workload, "unit": "bytes per launch", "schedule": {"engines": 1, "batch_bases": batch_bases},
           "copy_kernel": {"bytes_read": COPY_BYTES, "bytes_written": COPY_BYTES,
                           "dram_read_bytes": dram_rd(cp), "dram_write_bytes": dram_wr(cp),
                           "fetch_size_kib": avg(cp, "FETCH_SIZE"), "write_size_kib": avg(cp, "WRITE_SIZE"),
                           "dram_read_scale": f_rd, "dram_write_scale": f_wr,
                           "fetch_size_scale": f_fetch, "write_size_scale": f_write},
           "kernels": {}, "stages": {}}
    rows = []
    for k in sorted(per, key=lambda x: -(dram_rd(x) or 0) * ndisp(x)):
        e = {"dispatches": ndisp(k)}
        if dram_rd(k) is not None and f_rd and f_wr and dram_wr(k) is not None:
            e.update(dram_read_bytes=dram_rd(k), dram_write_bytes=dram_wr(k), ea_read_bytes=rd_bytes(k),
                     ea_write_bytes=wr_bytes(k), hbm_bytes=int(dram_rd(k) * f_rd + dram_wr(k) * f_wr))
        if avg(k, "FETCH_SIZE") is not None and f_fetch:
            e["fetch_size_bytes_scaled"] = int(avg(k, "FETCH_SIZE") * 1024 * f_fetch)
        if avg(k, "WRITE_SIZE") is not None and f_write:
            e["write_size_bytes_scaled"] = int(avg(k, "WRITE_SIZE") * 1024 * f_write)
        vi = avg(k, "SQ_INSTS_VALU_sum")
        if vi is not None:
            ns = sum(dur[k].values()) / len(dur[k])
            e.update(valu_insts=vi, valu_issue_frac=vi / (ns * 1e-9 * VALU_WAVE_INSTR_PEAK),
                     salu_insts=avg(k, "SQ_INSTS_SALU_sum"),
                     valu_int32_insts=avg(k, "SQ_INSTS_VALU_INT32_sum"))
            act, thr = avg(k, "SQ_ACTIVE_INST_VALU_sum"), avg(k, "SQ_THREAD_CYCLES_VALU_sum")
            if act:
                e["valu_active_lane_pct"] = 100.0 * thr / (act * 64) if thr is not None else None
        out["kernels"][k] = e
        rows.append((k, e))
    out["stages"] = pmc_stages(out["kernels"])
    lines = ["%-28s %6s %14s %14s %12s %12s %9s %7s" % ("kernel", "disp", "dram_rd_MB", "dram_wr_MB", "hbm_MB",
                                                        "fetchws_MB", "valu_iss", "lanes%")]
    for k, e in rows:
        f = lambda x, s=1e6: ("%.1f" % (e[x] / s)) if e.get(x) is not None else "-"
        lines.append("%-28s %6d %14s %14s %12s %12s %9s %7s" % (
            k[:28], e["dispatches"], f("dram_read_bytes"), f("dram_write_bytes"), f("hbm_bytes"),
            ("%.1f" % ((e["fetch_size_bytes_scaled"] + e["write_size_bytes_scaled"]) / 1e6))
            if "fetch_size_bytes_scaled" in e and "write_size_bytes_scaled" in e else "-",
            ("%.3f" % e["valu_issue_frac"]) if e.get("valu_issue_frac") is not None else "-",
            ("%.1f" % e["valu_active_lane_pct"]) if e.get("valu_active_lane_pct") is not None else "-"))
    return out, "\n".join(lines)


def schedule_note(trials):
    """How the schedule was chosen: the autotuner's trials, or fixed."""
    if not trials:
        return "fixed"
    n_seed = sum(1 for t in trials if "seed_lanes" in t)
    n_lanes = sum(1 for t in trials if "lookup_lanes" in t)
    n_sched = len(trials) - n_lanes - n_seed
    lanes = ("seeding form over %d lane counts, " % n_seed if n_seed else "") + \
            ("lookup stage form over %d lane counts, then " % n_lanes if n_lanes else "")
    return "autotuned: %s(pass, engines, batch) over %d triples" % (lanes, n_sched)


def pct(v, p):
    v = sorted(v)
    return v[min(len(v) - 1, int(math.ceil(p / 100.0 * len(v))) - 1)]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def kernel_report(ctx, kd, kd_steps, kd_source, run_batch, workload):
    """Roofline of the dominant kernel, lookup and extension figures from
    per-kernel (ms, launches, items, bytes, cells) deltas over kd_steps steps."""
    z = (0, 0, 0, 0, 0)
    # "fused_pass" times a whole recorded graph (every stage of a batch), not one kernel
    kd = {k: v for k, v in kd.items() if k != "fused_pass"}
    dom = max(kd, key=lambda k: kd[k][0])
    ms, launches, items, nbytes, _ = kd[dom]
    per_launch_ms = ms / max(1, launches)
    bytes_per_launch = nbytes / max(1, launches)
    achieved = bytes_per_launch / (per_launch_ms / 1000.0) / 1e9 if per_launch_ms > 0 else 0.0
    # HBM bytes per launch from the committed counter profile of this workload's
    # one-engine run (rocprofv3 --pmc, TCC read/write requests destined for DRAM,
    # copy-kernel calibrated; bench.py --summarize-pmc): carried from the profile,
    # not measured in this run
    tj = {}
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
    measured = tj.get("stages", {}) if tj.get("workload") == workload else {}
    traffic = measured[dom]["hbm_bytes_per_launch"] if dom in measured else None
    # the profile's batch size may differ from this run's schedule: per-launch
    # traffic then scales with the bases per batch (noted in the line)
    prof_batch = tj.get("schedule", {}).get("batch_bases")
    traffic_note = ("carried from profiles/pmc_traffic.json (rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B and "
                    "TCC_EA0_WRREQ_WRITE_DRAM_32B of the same one-engine command, k_hbm_copy calibrated), "
                    "not measured in this run")
    if traffic is not None and prof_batch and prof_batch != run_batch:
        traffic = int(traffic * run_batch / prof_batch)
        traffic_note += "; profiled at %d Mbases per batch, scaled to %d" % (prof_batch >> 20, run_batch >> 20)
    # index lookups: useful bytes (counted per probe) against the DRAM-side traffic
    lookup = None
    if kd.get("k_hits", z)[0] > 0:
        ms_, l_, _, b_, _ = kd["k_hits"]
        avg = ms_ / max(1, l_)
        lookup = {"kernel": "k_hits", "avg_launch_ms": round(avg, 3), "useful_bytes_per_launch": int(b_ / max(1, l_)),
                  "useful_gbps": round(b_ / (ms_ / 1000.0) / 1e9, 2), "peak_gbps": HBM_PEAK_GBS}
        if "k_hits" in measured:
            tb = measured["k_hits"]["hbm_bytes_per_launch"]
            if prof_batch and prof_batch != run_batch:
                tb = int(tb * run_batch / prof_batch)
            lookup.update({"traffic_bytes_per_launch": tb, "traffic_gbps": round(tb / (avg / 1000.0) / 1e9, 1),
                           "traffic_source": "carried from profiles/pmc_traffic.json"})
    # SSW score passes: DP cells computed (query rows x the columns each pass ran,
    # counted on the device) per second against the VALU integer peak (256 CUs x
    # 4 SIMDs x 16 lanes x 2.4 GHz = 39.3 T int32 ops/s; 10 ops per cell of the
    # striped recurrence: profile add, H from three maxima, the column maximum,
    # E and F each two subtractions and a maximum)
    def ssw_rate(name):
        ms_, _, _, _, cells = kd.get(name, z)
        if not cells or not ms_:
            return None
        gcups = cells / (ms_ / 1000.0) / 1e9
        return {"kernel": name, "cells_per_step": int(cells / kd_steps), "gcups": round(gcups, 1),
                "ops_per_cell": 10, "valu_int_peak_tops": VALU_INT_PEAK_TOPS,
                "frac": round(gcups * 10 / (VALU_INT_PEAK_TOPS * 1000.0), 4)}
    extension = ssw_rate("k_ssw_byte_fwd")
    if extension is not None:
        for key, name in (("word", "k_ssw_word_fwd"), ("byte_reverse", "k_ssw_byte_rev"),
                          ("word_reverse", "k_ssw_word_rev")):
            rate = ssw_rate(name)
            if rate is not None:
                extension[key] = rate
        # VALU issue of the score passes by their counters (SQ_INSTS_VALU over
        # the kernel's duration), from the committed profile of this workload
        if tj.get("workload") == workload:
            iss = {}
            for kname, e in tj.get("kernels", {}).items():
                if kname.startswith("k_ssw_pass<") and e.get("valu_issue_frac") is not None:
                    form = "byte" if ", true," in kname else "word"
                    if e["dispatches"] > iss.get(form, (0, 0))[0]:
                        iss[form] = (e["dispatches"], round(e["valu_issue_frac"], 3))
            if iss:
                extension["valu_issue_frac_from_profile"] = {k: v[1] for k, v in iss.items()}
    copy_gbps = ctx.hbm_copy_gbps()
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_note,
            "kernel": dom,
            "avg_launch_ms": round(per_launch_ms, 3), "bytes_per_launch": int(bytes_per_launch),
            "measured_copy_gbps": round(copy_gbps, 1),
            "kernel_ms_per_step": {k: round(v[0] / kd_steps, 2) for k, v in kd.items()},
            "kernel_gbps": {k: round(v[3] / (v[0] / 1000.0) / 1e9, 2) for k, v in kd.items() if v[3] and v[0]},
            "kernel_times_from": kd_source}
    if dom.startswith("k_ssw_"):
        # an SSW score pass is bound by the VALU, not by HBM (its jobs stay in
        # registers): cells x 10 int32 ops against the VALU integer peak, the
        # counter-measured instruction issue beside it
        ms_, launches_, _, _, cells = kd[dom]
        tops = cells * 10 / (ms_ / 1000.0) / 1e12 if ms_ and cells else 0.0
        roof.update({"bound": "valu", "unit": "TOP/s", "achieved": round(tops, 3), "peak": VALU_INT_PEAK_TOPS,
                     "frac": round(tops / VALU_INT_PEAK_TOPS, 5), "traffic": None,
                     "traffic_source": "an SSW score pass keeps its job in registers (counter traffic < 0.01 GB "
                                       "per launch)",
                     "bytes_per_launch": 0,
                     "cells_per_launch": int(cells / max(1, launches_)),
                     "valu_note": "10 int32 ops per DP cell; VALU issue by the counters in extension"})
    return roof, lookup, extension


def other_pass(ix, sa, reads, opts, threads, steps, n_units):
    """Times the batch pass the line did not use, on the same schedule (engines,
    batch size, lookup lanes): (units/s, the last step's stats)."""
    o = sa.options(threads=threads, slots=opts.slots, batch_bases=opts.batch_bases, lookup_lanes=opts.lookup_lanes,
                   seed_lanes=opts.seed_lanes, fused=not opts.fused)
    ix.map_reads(reads, o, keep_output=False)  # sizes (and records) the pass's buffers
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = ix.map_reads(reads, o, keep_output=False)
    return n_units * steps / (time.perf_counter() - t0), st


def pass_note(st, steps_stats=None):
    """What the fused pass did in a step (its batches, graph replays, speculative
    extensions made and used by the decisions)."""
    return {"fused_batches": st.fused_batches, "graph_replays": st.graph_replays,
            "speculative_extensions": st.spec_requests, "speculative_used": st.spec_used,
            "anchor_overflows": st.fused_overflows}


def replay_note(st):
    """The host replay of the decisions in a step: records decided per pass
    (committed or as probes) and thread time per record."""
    return {"records_committed": st.replay_committed, "records_probed": st.replay_probed,
            "probes_taken_over": st.replay_accepted,
            "us_per_committed": round(1000 * st.replay_commit_ms / max(1, st.replay_committed), 3),
            "us_per_probed": round(1000 * st.replay_probe_ms / max(1, st.replay_probed), 3),
            "us_writing_per_committed": round(1000 * st.replay_write_ms / max(1, st.replay_committed), 3),
            "thread_ms": round(st.replay_commit_ms + st.replay_probe_ms, 1)}


def fastq_mean_length(path, n=500):
    """Mean length of the first n records (the reference's read-length estimate,
    src/main.rs:1001-1012)."""
    tot = k = 0
    with open(path, "rb") as f:
        for i, line in enumerate(f):
            if i % 4 == 1:
                tot += len(line.rstrip(b"\r\n"))
                k += 1
                if k == n:
                    break
    return tot // max(1, k)


def bench_workload(a, sa, synth, cores, threads):
    """The other workloads, each on one resident index:

    human  -- 3.1 Gbp reference (24 contigs, seed 11, 45% of it copies of
              300 bp-50 kbp segments at 0-15% divergence), full index in HBM
              (bits 28, about 620M 16-byte entries); 2x250 pairs at 3%
              substitutions and 0.3% indels, mapped in alignment mode (SAM) and
              in mapping-only mode (-x, PAF).
    se_var -- the 100 Mbp reference; 2M single-end reads of 50-500 bp, 15% of
              them started inside copied repeats; batched (autotuned engines)
              against the plain one-engine path.
    The oracle leg maps a sample of the same reads on the host cores and its SAM
    is compared with the library's on that sample."""
    human = a.workload == "human"
    t0 = time.time()
    if human:
        w = synth.make_human(a.data_dir, a.pairs)
    else:
        w = synth.make("se_var", a.data_dir, a.pairs if a.pairs != 1_000_000 else 2_000_000)
    gen_s = time.time() - t0
    print("data ready in %.1fs: %s" % (gen_s, w), file=sys.stderr, flush=True)
    read_len = 250 if human else fastq_mean_length(w["reads1"])
    ctx = sa.Context(0)
    t0 = time.time()
    ix = sa.Index(ctx, w["ref"], read_length=read_len, threads=threads, device_build=not a.host_index)
    index_s = time.time() - t0
    params, n_entries, n_contigs = ix.params()
    print("index %.1fs: %d entries, bits %d, %d contigs, read length %d" %
          (index_s, n_entries, params.bits, n_contigs, read_len), file=sys.stderr, flush=True)
    reads = sa.Reads(w["reads1"], w.get("reads2"))
    t0 = time.perf_counter()
    reads.upload(ix)
    upload_s = time.perf_counter() - t0
    n_units = len(reads)
    unit = "read pairs/s" if human else "reads/s"
    opts = sa.options(threads=threads, slots=a.slots, batch_bases=a.batch_bases, lookup_lanes=a.lookup_lanes,
                      seed_lanes=a.seed_lanes)
    trials = None
    if a.autotune and not a.slots and not a.batch_bases:
        t0 = time.perf_counter()
        trials = ix.autotune(reads, opts, sample_pairs=a.tune_pairs or min(n_units, 500_000))
        print("autotune: %d lookup lanes, %d engines, %d Mbases per batch (%d trials, %.1f s)" %
              (opts.lookup_lanes, opts.slots, opts.batch_bases >> 20, len(trials), time.perf_counter() - t0),
              file=sys.stderr, flush=True)
    if a.pass_ != "auto":
        opts.fused = int(a.pass_ == "fused")
    runs = [("sam", 0, opts.slots), ("paf", 1, opts.slots)] if human else [("sam", 0, opts.slots)]
    if not human and a.single_steps > 0:
        runs.append(("plain", 0, 1))
    res = {}
    for name, mode, slots in runs:
        o = sa.options(mode=mode, threads=threads, slots=slots, batch_bases=a.batch_bases if slots == 1 else opts.batch_bases,
                       lookup_lanes=opts.lookup_lanes, seed_lanes=opts.seed_lanes,
                       fused=bool(opts.fused) and name == "sam")
        steps = a.steps if name != "plain" else a.single_steps
        for i in range(a.warmup if name != "plain" else 1):
            _, st = ix.map_reads(reads, o, keep_output=False)
            print("%s warmup %d: %.1f ms" % (name, i, st.total_ms), file=sys.stderr, flush=True)
        k0 = ix.kernel_times()
        t0 = time.perf_counter()
        for i in range(steps):
            _, st = ix.map_reads(reads, o, keep_output=False)
            print("%s step %d: %.1f ms (rounds %d, extension requests %d)" %
                  (name, i, st.total_ms, st.rounds, st.requests), file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        k1 = ix.kernel_times()
        z = (0, 0, 0, 0, 0)
        kd = {k: tuple(k1[k][x] - k0.get(k, z)[x] for x in range(5)) for k in k1}
        res[name] = (n_units * steps / wall, wall, st, kd, steps)
    value, wall, st, kd, _ = res["sam"]
    other, other_st = other_pass(ix, sa, reads, opts, threads, max(1, min(a.steps, 3)), n_units)
    fused_rate, plain_rate = (value, other) if opts.fused else (other, value)
    print("fused pass %.0f, plain pass %.0f %s (same schedule)" % (fused_rate, plain_rate, unit), file=sys.stderr,
          flush=True)
    if "plain" in res:  # per-kernel figures from the one-engine pass (kernels alone on the device)
        _, _, _, kdp, sp = res["plain"]
        roof, lookup, extension = kernel_report(ctx, kdp, sp, "the one-engine pass (%d steps)" % sp,
                                                a.batch_bases or (64 << 20), a.workload)
    else:
        roof, lookup, extension = kernel_report(ctx, kd, a.steps, "the %d-engine SAM steps" % (opts.slots or 4),
                                                opts.batch_bases or (64 << 20), a.workload)
    cpu = None
    parity = None
    if not a.no_cpu and a.cpu_pairs != 0:
        # enough records to keep all host threads busy (one chunk per thread and more)
        sample = min(a.cpu_pairs or (100_000 if human else 200_000), n_units)
        osam = os.path.join(a.data_dir, "oracle_sample.sam")
        try:
            v, t, (s1, s2) = cpu_baseline(w["ref"], w["reads1"], w.get("reads2"), sample, cores, a.data_dir,
                                          out=osam, t1_chunks=1, read_length=read_len)
            cpu = {"value": round(v, 1), "unit": unit, "cores": cores, "kind": "port",
                   "sample": "first %d %s of the same workload, oracle -t %d, SAM; index build excluded "
                             "(%.0f s, one thread)" % (sample, "pairs" if human else "reads", cores, t["index_s"]),
                   "label": "CPU restatement of reference v0.18.0-alpha (oracle/), not the upstream binary",
                   "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                   "single_thread_value": round(t["t1_records"] / t["t1_s"], 1) if t.get("t1_s") else None,
                   "stage_seconds_summed_over_threads": t.get("stages_s")}
            mine = ix.map_files(s1, s2, threads=threads, fused=bool(opts.fused), slots=opts.slots,
                                batch_bases=opts.batch_bases, lookup_lanes=opts.lookup_lanes,
                                seed_lanes=opts.seed_lanes)
            with open(osam, "rb") as f:
                theirs = f.read()
            parity = {"records": sample, "identical_sam": mine == theirs, "bytes": len(theirs),
                      "compared": "library SAM (the measured pass and schedule) vs oracle SAM (--no-PG) of the "
                                  "first %d records" % sample}
            print("oracle sample: %.0f %s; SAM identical: %s" % (v, unit, parity["identical_sam"]), file=sys.stderr,
                  flush=True)
        except Exception as e:  # reported, never fatal
            cpu = {"value": None, "unit": unit, "cores": cores, "kind": "port", "sample": "failed: %s" % e}
    if human:
        metric = "read pairs/sec at 2x250 on 3.1 Gbp ref (SAM; -x PAF beside it)"
        desc = "3.1 Gbp reference (24 contigs, 45%% repeats), %d pairs 2x250, 3%% subs, 0.3%% indels" % n_units
        data = "synthetic (strobealign_amd/synth.py human, seed 11)"
    else:
        metric = "reads/sec, single-end 50-500 bp on 100 Mbp ref (batched vs plain per-batch)"
        desc = "100 Mbp reference (20x5 Mbp, 30%% repeats), %d SE reads of 50-500 bp, 15%% from repeats" % n_units
        data = "synthetic (strobealign_amd/synth.py se_var, seed 71)"
    config = {"workload": desc, "records": n_units, "read_len": read_len, "host_threads": threads,
              "index_entries": n_entries, "index_bits": params.bits, "index_build_s": round(index_s, 1),
              "index_build": "host, %d threads" % threads if a.host_index else "device (FASTA read on the host)",
              "data_generation_s": round(gen_s, 1), "read_upload_ms": round(1000 * upload_s, 2),
              "extension_requests_per_step": st.requests, "rounds_per_step": st.rounds,
              "engines": opts.slots or 4, "batch_bases": opts.batch_bases or (64 << 20),
              "lookup_lanes": opts.lookup_lanes or 16, "seed_lanes": opts.seed_lanes or 64,
              "schedule": schedule_note(trials),
              "pass": "fused" if opts.fused else "plain",
              "fused_pass_per_s": round(fused_rate, 1), "plain_pass_per_s": round(plain_rate, 1),
              "fused_over_plain_pass": round(fused_rate / plain_rate, 3),
              "fused_pass_step": pass_note(st if opts.fused else other_st),
              "host_replay_step": replay_note(st)}
    if human:
        config["paf_pairs_per_s"] = round(res["paf"][0], 1)
    if "plain" in res:
        config["plain_one_engine_per_s"] = round(res["plain"][0], 1)
        config["batched_over_plain"] = round(value / res["plain"][0], 3)
    out = {
        "metric": metric,
        "value": round(value, 1),
        "unit": unit,
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1000.0 * wall / a.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": data,
        "config": config,
        "roofline": roof,
        "lookup": lookup,
        "extension": extension,
        "cpu_baseline": cpu,
        "parity_sample": parity,
    }
    if human:
        out["paf_kernel_ms_per_step"] = {k: round(v[0] / a.steps, 2) for k, v in res["paf"][3].items() if v[0]}
    print(json.dumps(out), flush=True)
    reads.close()
    ix.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["pe100m", "human", "se_var"], default="pe100m",
                    help="pe100m: 100 Mbp, 2x150 (the headline line); human: 3.1 Gbp, 2x250, SAM and -x; "
                         "se_var: 100 Mbp, 2M single-end 50-500 bp, batched vs plain")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--threads", type=int, default=0, help="host threads (default: all cores)")
    ap.add_argument("--data-dir", default=os.environ.get("SA_BENCH_DATA", "/tmp/strobealign_amd_bench"))
    ap.add_argument("--cpu-pairs", type=int, default=None,
                    help="oracle sample size (0: skip; default 200,000, human 100,000)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--host-index", action="store_true",
                    help="human/se_var: build the index on the host instead of the device")
    ap.add_argument("--slots", type=int, default=0, help="device batches in flight (0: library default)")
    ap.add_argument("--batch-bases", type=int, default=0, help="bases per device batch (0: library default)")
    ap.add_argument("--pass", dest="pass_", choices=["auto", "fused", "plain"], default="auto",
                    help="batch pass: fused (seeding to the expected extensions as one replayed graph per batch), "
                         "plain (per-batch stages, extensions requested by the decisions), or auto (autotuner)")
    ap.add_argument("--seed-lanes", type=int, default=0, choices=[0, 1, 64],
                    help="lanes per read in seeding: 1 (a thread) or 64 (a wave); 0: autotuner or library default")
    ap.add_argument("--lookup-lanes", type=int, default=0, choices=[0, 1, 8, 16, 64],
                    help="lanes per read orientation in the lookup stage (0: autotuned, else library default)")
    ap.add_argument("--single-steps", type=int, default=2, help="timed steps of the one-engine comparison (0: skip)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="use the library's default schedule instead of sa_autotune")
    ap.add_argument("--tune-pairs", type=int, default=0, help="pairs the autotuner times (0: the whole workload)")
    ap.add_argument("--serve-requests", type=int, default=400, help="requests of the serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-varied", type=int, default=400,
                    help="requests of a serving run with 100-300 bp reads on an r250 index (0: skip)")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="DB",
                    help="write profiles/pmc_traffic.json from rocprofv3 --pmc databases of this workload's "
                         "one-engine run (one counter pass each) and print the per-kernel table; no GPU use")
    a = ap.parse_args()
    if a.summarize_pmc:
        res, table = summarize_pmc(a.summarize_pmc, a.workload, a.batch_bases or (64 << 20))
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json"), "w") as f:
            json.dump(res, f, indent=1)
        print(table)
        sys.exit(0)
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import strobealign_amd as sa
    from strobealign_amd import synth

    cores = host_cores()
    threads = a.threads or cores
    if a.workload != "pe100m":
        return bench_workload(a, sa, synth, cores, threads)
    if a.cpu_pairs is None:
        a.cpu_pairs = 200_000
    t0 = time.time()
    w = synth.make("pe100m", a.data_dir, a.pairs)
    gen_s = time.time() - t0
    print("data ready in %.1fs: %s" % (gen_s, w), file=sys.stderr, flush=True)
    ctx = sa.Context(0)
    t0 = time.time()
    ix = sa.Index(ctx, w["ref"], read_length=150, threads=threads)
    index_s = time.time() - t0
    t0 = time.perf_counter()
    reads = sa.Reads(w["reads1"], w["reads2"])
    host_parse_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    reads.upload(ix)  # inputs resident in HBM before the timed region
    upload_s = time.perf_counter() - t0
    # the same files parsed on the device (sa_reads_load_device): file read,
    # copy in, newline scan, record checks, staging; records must be equal
    t0 = time.perf_counter()
    dreads = sa.Reads(w["reads1"], w["reads2"], device=ix)
    dev_ingest_s = time.perf_counter() - t0
    ingest = {"host_parse_ms": round(1000 * host_parse_s, 1), "host_upload_ms": round(1000 * upload_s, 1),
              "device_ingest_wall_ms": round(1000 * dev_ingest_s, 1),
              "device_parse_ms": round(dreads.ingest_ms, 2),
              "records_equal_sampled": all(dreads.record(i, m) == reads.record(i, m)
                                           for i in range(0, len(reads), max(1, len(reads) // 1000)) for m in (0, 1))}
    dreads.close()
    print("ingest: host parse %.0f ms + upload %.0f ms; device ingest %.0f ms wall (%.1f ms on the device)" %
          (1000 * host_parse_s, 1000 * upload_s, 1000 * dev_ingest_s, ingest["device_parse_ms"]),
          file=sys.stderr, flush=True)
    opts = sa.options(threads=threads, slots=a.slots, batch_bases=a.batch_bases, lookup_lanes=a.lookup_lanes,
                      seed_lanes=a.seed_lanes)
    trials = None
    if a.autotune and not a.slots and not a.batch_bases:
        # schedule (engines x batch size) chosen by timing the workload itself
        # (sa_autotune; outputs are identical for every schedule)
        t0 = time.perf_counter()
        trials = ix.autotune(reads, opts, sample_pairs=a.tune_pairs)
        print("autotune: %s pass, %d lookup lanes, %d engines, %d Mbases per batch (%d trials, %.1f s)" %
              ("fused" if opts.fused else "plain", opts.lookup_lanes, opts.slots, opts.batch_bases >> 20,
               len(trials), time.perf_counter() - t0), file=sys.stderr, flush=True)
    if a.pass_ != "auto":
        opts.fused = int(a.pass_ == "fused")
    print("index %.1fs, %d pairs loaded, staged on the device in %.1f ms" % (index_s, len(reads), 1000 * upload_s),
          file=sys.stderr, flush=True)
    for i in range(a.warmup):
        _, st = ix.map_reads(reads, opts, keep_output=False)
        print("warmup %d: %.1f ms" % (i, st.total_ms), file=sys.stderr, flush=True)
    k0 = ix.kernel_times()
    stats = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        _, st = ix.map_reads(reads, opts, keep_output=False)
        stats.append(st)
        print("step %d: %.1f ms (summed over batches: nams %.1f, extend %.1f, host %.1f; rounds %d)" %
              (i, st.total_ms, st.nam_ms, st.extend_ms, st.host_ms, st.rounds), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    k1 = ix.kernel_times()
    n_pairs = len(reads)
    value = n_pairs * a.steps / wall
    other, other_st = other_pass(ix, sa, reads, opts, threads, max(1, min(a.steps, 3)), n_pairs)
    fused_rate, plain_rate = (value, other) if opts.fused else (other, value)
    print("fused pass %.0f, plain pass %.0f pairs/s (same schedule)" % (fused_rate, plain_rate), file=sys.stderr,
          flush=True)
    # the plain per-batch drop-in beside it: one engine, one batch at a time (as a
    # caller handing batches to sa_map_reads one after another would run it)
    single = None
    if a.single_steps > 0:
        o1 = sa.options(threads=threads, slots=1, batch_bases=a.batch_bases, lookup_lanes=opts.lookup_lanes,
                        seed_lanes=opts.seed_lanes)
        ix.map_reads(reads, o1, keep_output=False)  # grows the engine's buffers to this schedule
        k1b = ix.kernel_times()
        t1 = time.perf_counter()
        for _ in range(a.single_steps):
            ix.map_reads(reads, o1, keep_output=False)
        single = n_pairs * a.single_steps / (time.perf_counter() - t1)
        print("one engine: %.0f pairs/s" % single, file=sys.stderr, flush=True)
    k2 = ix.kernel_times()
    # dominant kernel over the timed steps: HIP-event time on the engines' streams;
    # algorithmic bytes are counted by the engine per launch (bytes each stage must
    # read and write: read bases and randstrobes, bucket words and entries probed,
    # hits, NAMs and chain anchors)
    z = (0, 0, 0, 0, 0)
    delta = {k: tuple(k1[k][x] - k0.get(k, z)[x] for x in range(5)) for k in k1}
    # Per-kernel figures come from the one-engine pass when it ran: there each kernel
    # has the device to itself, so its HIP-event time is its own duration (with
    # several engines, events also count the time a kernel shares the device or
    # waits behind another engine's kernels).
    kd = delta
    kd_steps = a.steps
    kd_source = "the %d-engine timed steps" % (opts.slots or 4)
    if single:
        kd = {k: tuple(k2[k][x] - k1b.get(k, z)[x] for x in range(5)) for k in k2}
        kd_steps = a.single_steps
        kd_source = "the one-engine pass (%d steps, %d Mbases per batch)" % (a.single_steps,
                                                                           (a.batch_bases or (64 << 20)) >> 20)
    run_batch = (a.batch_bases or (64 << 20)) if single else (opts.batch_bases or (64 << 20))
    roof, lookup, extension = kernel_report(ctx, kd, kd_steps, kd_source, run_batch, "pe100m")
    serving = None
    if a.serve_requests > 0:
        # request loop over the same resident index: Poisson arrivals, then the
        # same requests submitted at once (saturation throughput)
        plan = request_plan(len(reads), a.serve_requests, a.serve_rate, 11)
        slots = a.slots or 4
        lat, swall, sst, (lag, sub_delay) = serve(sa, ix, reads, threads, slots, plan, True)
        pairs = sum(n for _, _, n in plan)
        _, sat_wall, sat_st, _ = serve(sa, ix, reads, threads, slots, plan, False)
        serving = {"requests": len(plan), "arrivals": "Poisson %.0f/s" % a.serve_rate,
                   "pairs_per_request": "log-uniform [1, 50000] (mean %.0f)" % (pairs / len(plan)),
                   "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
                   "mean_ms": round(sum(lat) / len(lat), 2), "max_ms": round(max(lat), 2),
                   "offered_pairs_per_s": round(pairs / plan[-1][0], 1),
                   "served_pairs_per_s": round(pairs / swall, 1),
                   "saturated_pairs_per_s": round(pairs / sat_wall, 1),
                   "saturated_clients": 4,
                   "batches": sst["batches"], "max_requests_per_batch": sst["max_requests_per_batch"],
                   "mean_requests_per_batch": round(sst["batch_requests"] / max(1, sst["batches"]), 2),
                   "arena_peak_bytes": sst["arena_peak"], "driver_max_lag_ms": round(1000 * lag, 2),
                   "max_submit_delay_ms": round(1000 * sub_delay, 2),
                   "latency_from": "scheduled Poisson arrival to the last output byte", "pass": "plain",
                   "engines": slots}
        print("serving: p50 %.1f ms, p99 %.1f ms, saturated %.0f pairs/s" %
              (serving["p50_ms"], serving["p99_ms"], serving["saturated_pairs_per_s"]), file=sys.stderr, flush=True)
    serving_varied = serve_varied(a, sa, synth, ctx, threads, a.slots or 4) if a.serve_varied > 0 else None
    cpu = None
    parity = None
    if not a.no_cpu and a.cpu_pairs > 0:
        try:
            sample = min(a.cpu_pairs, n_pairs)
            osam = os.path.join(a.data_dir, "oracle_pe100m_sample.sam")
            v, t, (s1, s2) = cpu_baseline(w["ref"], w["reads1"], w["reads2"], sample, cores, a.data_dir, out=osam)
            # the headline claims bit-exact SAM: the library maps the same sample
            # records and its SAM is compared byte for byte with the oracle's
            mine = ix.map_files(s1, s2, threads=threads, fused=bool(opts.fused), slots=opts.slots,
                                batch_bases=opts.batch_bases, lookup_lanes=opts.lookup_lanes,
                                seed_lanes=opts.seed_lanes)
            with open(osam, "rb") as f:
                theirs = f.read()
            parity = {"records": sample, "identical_sam": mine == theirs, "bytes": len(theirs),
                      "compared": "library SAM (the measured pass and schedule) vs oracle SAM (--no-PG) of the "
                                  "first %d pairs" % sample}
            print("oracle sample: %.0f pairs/s; SAM identical: %s" % (v, parity["identical_sam"]), file=sys.stderr,
                  flush=True)
            cpu = {"value": round(v, 1), "unit": "read pairs/s", "cores": cores, "kind": "port",
                   "sample": "first %d pairs of the same workload, oracle -t %d, index build excluded (%.1f s map)"
                             % (min(a.cpu_pairs, n_pairs), cores, t["map_s"]),
                   "label": "CPU restatement of reference v0.18.0-alpha (oracle/), not the upstream binary",
                   "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                   "single_thread_value": round(t["t1_records"] / t["t1_s"], 1) if t.get("t1_s") else None,
                   "single_thread_sample": "first %d pairs on one thread" % t.get("t1_records", 0),
                   "stage_seconds_summed_over_threads": t.get("stages_s")}
        except Exception as e:  # reported, never fatal
            cpu = {"value": None, "unit": "read pairs/s", "cores": cores, "kind": "port", "sample": "failed: %s" % e}
    out = {
        "metric": "read pairs/sec at 2x150 on 100 Mbp ref, bit-exact SAM; p50/p99 latency",
        "value": round(value, 1),
        "unit": "read pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1000.0 * wall / a.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (strobealign_amd/synth.py pe100m, seed 7)",
        "config": {"workload": "100 Mbp reference (20x5 Mbp, 30%% repeats), %d pairs 2x150, SAM output" % n_pairs,
                   "pairs": n_pairs, "read_len": 150, "host_threads": threads,
                   "read_upload_ms": round(1000 * upload_s, 2),
                   "pairs_per_s_incl_pcie": round(n_pairs * a.steps / (wall + a.steps * upload_s), 1),
                   "pairs_per_s_incl_device_ingest": round(n_pairs * a.steps / (wall + a.steps * dev_ingest_s), 1),
                   "ingest": ingest,
                   "rounds_per_step": stats[-1].rounds, "extension_requests_per_step": stats[-1].requests,
                   "engines": opts.slots or 4, "batch_bases": opts.batch_bases or (64 << 20),
                   "lookup_lanes": opts.lookup_lanes or 16, "seed_lanes": opts.seed_lanes or 64,
                   "schedule": schedule_note(trials),
                   "pass": "fused" if opts.fused else "plain",
                   "fused_pass_pairs_per_s": round(fused_rate, 1), "plain_pass_pairs_per_s": round(plain_rate, 1),
                   "fused_over_plain_pass": round(fused_rate / plain_rate, 3),
                   "fused_pass_step": pass_note(stats[-1] if opts.fused else other_st),
                   "host_replay_step": replay_note(stats[-1]),
                   "plain_one_engine_pairs_per_s": round(single, 1) if single else None,
                   "batched_over_plain": round(value / single, 3) if single else None},
        # request latency of the serving configuration as specified (one read length
        # per request, uniform in [100, 300]); the 2x150 serving run is beside it
        "latency_ms": ({"p50": serving_varied["p50_ms"], "p99": serving_varied["p99_ms"],
                        "from": "serving_varied_lengths"} if serving_varied else
                       {"p50": serving["p50_ms"], "p99": serving["p99_ms"], "from": "serving"} if serving else None),
        "serving": serving,
        "serving_varied_lengths": serving_varied,
        "roofline": roof,
        "lookup": lookup,
        "extension": extension,
        "cpu_baseline": cpu,
        "parity_sample": parity,
    }
    if parity is not None and not parity["identical_sam"]:
        out["metric"] += " (SAM sample NOT identical to the oracle)"
    print(json.dumps(out), flush=True)
    reads.close()
    ix.close()
    ctx.close()


if __name__ == "__main__":
    main()
