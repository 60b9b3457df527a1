"""Size-class autotuner for batched joins (fij_join_tuning).

For every size class (ceil(log2 |R|), ceil(log2 |S|)) of the batched
workload it times, on the current GPU, a grid of configurations of batched
joins of that class and keeps the fastest:

* the small-join path (one block per join, LDS table), where the class
  qualifies for it;
* the two-pass plan with total radix bits b0 - 1, b0, b0 + 1 (b0 from the
  closed-form rule), each split three ways between the passes, a pass-1 line
  buffer of the deepest ring that fits LDS or of at most 4 lines per
  partition, and pass-1 blocks of 512 or 1024 threads.

Every candidate's per-join results must equal the closed form of the
workload (every probe key matches once; R.value = key index + 1 is not
assumed: the check is count = |S| and SUM(S.value) = |S|(|S|-1)/2) and the
results of the rule's configuration, bit for bit.  A candidate replaces the
rule only when it is at least 3 % faster (median of three timed
submissions, HIP events around the graph launch).

    python fast-interconnects_amd/autotune.py [--out-json PATH] [--out-inc PATH]
        [--min-log2r 10] [--max-log2r 22] [--ratios 2,3,4] [--tuples-log2 26]

Writes the measurements (JSON) and the table the library compiles in
(csrc/tuning_table.inc).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import fij  # noqa: E402

SEED = 1234
SMALL_BUILD_I32 = 12288
SMALL_PROBE = 1 << 22


def log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def candidates(rule, build_len, probe_len):
    out = []
    if build_len <= SMALL_BUILD_I32 and probe_len <= SMALL_PROBE:
        out.append({"small": 1, "radix_bits_1": rule["radix_bits_1"], "radix_bits_2": rule["radix_bits_2"],
                    "ring_lines": 0, "block": 1024})
    b0 = rule["radix_bits_1"] + rule["radix_bits_2"]
    seen = set()
    for b in (b0 - 1, b0, b0 + 1):
        if b < 2:
            continue
        mid = (b + 1) // 2
        for b1 in (mid - 1, mid, mid + 1):
            b2 = b - b1
            if not (1 <= b1 <= 9 and 1 <= b2 <= 12) or (b1, b2) in seen:
                continue
            seen.add((b1, b2))
            for ring in (0, 4):
                for block in (512, 1024):
                    out.append({"small": 0, "radix_bits_1": b1, "radix_bits_2": b2, "ring_lines": ring,
                                "block": block})
    return out


def make_joins(n_r, n_s, count, dtype):
    words = count * 2 * (n_r + n_s)
    arena = fij.DeviceArray(words, dtype)
    joins, off = [], 0
    for i in range(count):
        rk = arena.view(off, off + n_r); off += n_r
        rp = arena.view(off, off + n_r); off += n_r
        sk = arena.view(off, off + n_s); off += n_s
        sp = arena.view(off, off + n_s); off += n_s
        fij.gen_permutation(rk, SEED + i, 1)
        fij.gen_iota(rp, 1)
        fij.gen_uniform(sk, SEED + i, 1, 1, n_r)
        fij.gen_iota(sp, 0)
        joins.append((rk, rp, sk, sp))
    fij.synchronize()
    return arena, joins


def time_config(dtype, cls, cfg, joins, reps=3):
    entry = dict(cfg, build_log2=cls[0], probe_log2=cls[1])
    fij.tuning_set(dtype, [entry])
    batch = fij.JoinBatch(dtype, len(joins), use_graph=True)
    got, _ = batch.execute(joins)  # warm-up (allocates the class's buffers)
    ms = []
    for _ in range(reps):
        again, st = batch.execute(joins)
        assert again == got, "repeated submissions differ"
        ms.append(st["gpu_ms"])
    del batch
    return float(np.median(ms)), got


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dtype", choices=["int32", "int64"], default="int32")
    ap.add_argument("--min-log2r", type=int, default=10)
    ap.add_argument("--max-log2r", type=int, default=22)
    ap.add_argument("--ratios", default="2,3,4", help="log2(|S| / |R|) classes")
    ap.add_argument("--tuples-log2", type=int, default=26, help="tuples per timed submission (log2)")
    ap.add_argument("--margin", type=float, default=0.03, help="required gain over the rule")
    ap.add_argument("--out-json", default=os.path.join(HERE, "tuning", "gfx950_int32_batched.json"))
    ap.add_argument("--out-inc", default=os.path.join(HERE, "csrc", "tuning_table.inc"))
    a = ap.parse_args()
    dtype = np.dtype(a.dtype).type
    ratios = [int(x) for x in a.ratios.split(",")]
    fij.tuning_set(dtype, [])  # the rule, for the reference configuration
    limits = fij.device_limits()
    report = {"gpu": limits.get("arch_name", ""), "compute_units": limits.get("compute_units", 0),
              "dtype": a.dtype, "tuples_per_submission": 1 << a.tuples_log2, "margin": a.margin, "classes": []}
    table = []
    t_start = time.time()
    for c in range(a.min_log2r, a.max_log2r + 1):
        for r in ratios:
            n_r = (3 << c) // 4
            n_s = (3 << (c + r)) // 4
            cls = (c, c + r)
            count = int(max(4, min(4096, (1 << a.tuples_log2) // (n_r + n_s))))
            fij.tuning_set(dtype, [])
            rule = fij.tuning_lookup(dtype, n_r, n_s)
            assert (rule["build_log2"], rule["probe_log2"]) == cls
            arena, joins = make_joins(n_r, n_s, count, dtype)
            want = [(n_s, None, n_s * (n_s - 1) // 2)] * count
            rule_cfg = {k: rule[k] for k in ("small", "radix_bits_1", "radix_bits_2", "ring_lines", "block")}
            rule_ms, rule_res = time_config(dtype, cls, rule_cfg, joins)
            for g, w in zip(rule_res, want):
                assert g[0] == w[0] and g[2] == w[2], f"class {cls}: rule result {g}"
            rows = []
            for cfg in candidates(rule, n_r, n_s):
                ms, res = time_config(dtype, cls, cfg, joins)
                assert res == rule_res, f"class {cls} config {cfg}: results differ from the rule's"
                rows.append(dict(cfg, gpu_ms=ms))
            best = min(rows, key=lambda x: x["gpu_ms"])
            chosen = best if best["gpu_ms"] < rule_ms * (1 - a.margin) else dict(rule_cfg, gpu_ms=rule_ms)
            tuples = count * (n_r + n_s)
            report["classes"].append({"build_log2": cls[0], "probe_log2": cls[1], "joins": count, "build_len": n_r,
                                      "probe_len": n_s, "rule": dict(rule_cfg, gpu_ms=rule_ms), "chosen": chosen,
                                      "candidates": rows,
                                      "chosen_gtuples_per_s": tuples / (chosen["gpu_ms"] / 1e3) / 1e9,
                                      "rule_gtuples_per_s": tuples / (rule_ms / 1e3) / 1e9})
            table.append(dict(chosen, build_log2=cls[0], probe_log2=cls[1], rule_ms=rule_ms))
            log(f"class {cls}: {count} joins of {n_r} x {n_s}; rule {rule_cfg['radix_bits_1']}+"
                f"{rule_cfg['radix_bits_2']} small={rule_cfg['small']} {rule_ms:.3f} ms; chosen "
                f"{chosen['radix_bits_1']}+{chosen['radix_bits_2']} small={chosen['small']} ring={chosen['ring_lines']} "
                f"block={chosen['block']} {chosen['gpu_ms']:.3f} ms ({len(rows)} candidates, "
                f"{time.time() - t_start:.0f} s)")
            del arena, joins
    fij.tuning_set(dtype, builtin=True)
    os.makedirs(os.path.dirname(a.out_json), exist_ok=True)
    with open(a.out_json, "w") as f:
        json.dump(report, f, indent=1)
    write_inc(a.out_inc, a.dtype, table, report)
    log(f"wrote {a.out_json} and {a.out_inc}")


def write_inc(path, dtype_name, table, report):
    """Keeps the other key type's list of an existing table."""
    other = {"int32": "kTunedI64", "int64": "kTunedI32"}[dtype_name]
    mine = {"int32": "kTunedI32", "int64": "kTunedI64"}[dtype_name]
    keep = "static const fij_join_tuning %s[] = {\n    {0, 0, 0, 0, 0, 0, 0, 0},\n};\n" % other
    if os.path.exists(path):
        text = open(path).read()
        start = text.find("static const fij_join_tuning %s[]" % other)
        if start >= 0:
            keep = text[start:text.index("};", start) + 3]
    lines = [
        "// Size-class tuning table (see fij_join_tuning in include/fij.h).",
        "// Written by fast-interconnects_amd/autotune.py; entries:",
        "// {build_log2, probe_log2, small, radix_bits_1, radix_bits_2, ring_lines, block, 0}.",
        f"// {mine}: measured on {report['gpu']} ({report['compute_units']} CUs), batched {dtype_name} joins,",
        f"// {report['tuples_per_submission']} tuples per timed submission; chosen / rule GPU ms per submission.",
    ]
    body = [f"static const fij_join_tuning {mine}[] = {{"]
    for e in table:
        body.append(f"    {{{e['build_log2']}, {e['probe_log2']}, {e['small']}, {e['radix_bits_1']}, "
                    f"{e['radix_bits_2']}, {e['ring_lines']}, {e['block']}, 0}},  "
                    f"// {e['gpu_ms']:.3f} / {e['rule_ms']:.3f} ms")
    body.append("    {0, 0, 0, 0, 0, 0, 0, 0},")
    body.append("};")
    blocks = [keep.rstrip() + "\n", "\n".join(body) + "\n"]
    if mine == "kTunedI32":
        blocks.reverse()
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n" + "".join(blocks))


if __name__ == "__main__":
    main()
