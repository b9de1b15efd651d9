"""Write tests/golden/stage0_runs.json: what the reference `Tree` prints about
its Stage-0 plan on generated inputs.

Run in the build container only (it needs oracle/_ref/Tree, compiled from the
reference sources by oracle/Makefile):

    python tests/golden/make_stage0.py

Stage 0 of the reference (src/main.cpp:140-189) samples the FASTQ file
(FastqPreReader::pre_read, src/FastqPreReader.h:128-300), histograms the root
prefixes of the sampled keys, clamps them to >= 1000, assigns them to the
classifiers by LPT and sizes the Bloom filters.  Tree prints three results of
it: "Average prefix count" (sum of the clamped counts / 256, :175-186), one
"Classifier i total prefix count" line per classifier (the LPT loads, :77-82)
and "Max Bloom Filter capacity" (:106).  These pin the oracle's restatement of
the whole Stage-0 chain, including the sampler's part boundaries.

Inputs: the deterministic generator (kmertreecounter_amd.generate), written
plain or gzip-compressed.  Cases cover a sample of three separate parts (files
above 6 MiB), overlapping 2 MiB reads (files below 6 MiB), gzip input (2 MiB
reads of the decompressed stream, up to the compressed size), k = 31 and 63,
several thread counts (classifier counts 1..6).  The reference's Stage-0
threads can race (a consumer can stop before the last block); a case is kept
when two runs print the same lines.
"""
from __future__ import annotations

import gzip
import json
import re
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import kmertreecounter_amd as K  # noqa: E402

REF = ROOT / "oracle" / "_ref"
OUT = Path(__file__).resolve().parent / "stage0_runs.json"

# (name, workload, scale, seed, k, n_thread, gzip)
CASES = [
    ("c1_s10_k31_t8", 1, 10.0, 11, 31, 8, False),
    ("c1_s10_k63_t16", 1, 10.0, 15, 63, 16, False),
    ("c1_s40_k31_t8", 1, 40.0, 21, 31, 8, False),
    ("c1_s100_k17_t6", 1, 100.0, 22, 17, 6, False),
    ("c1_s10_k31_t12_gz", 1, 10.0, 11, 31, 12, True),
    ("c2_s200_k31_t8", 2, 200.0, 2, 31, 8, False),
    ("c3_s200_k31_t10", 3, 200.0, 3, 31, 10, False),
    ("c3_s50_k31_t8_gz", 3, 50.0, 7, 31, 8, True),
    ("c1_full_k31_t8", 1, 1.0, 1, 31, 8, False),
]


def stage0_lines(text: str) -> dict:
    avg = re.search(r"Average prefix count: (\d+)", text)
    cap = re.search(r"Max Bloom Filter capacity: (\d+)", text)
    loads = [(int(i), int(c)) for i, c in re.findall(r"Classifier (\d+) total prefix count: (\d+)", text)]
    if not avg or not cap or not loads:
        raise RuntimeError("Stage-0 lines missing from the reference output")
    return {"average": int(avg.group(1)), "max_bloom_capacity": int(cap.group(1)),
            "loads": sorted(c for _, c in loads), "classifiers": len(loads)}


def run_case(name, cfg, scale, seed, k, n_thread, gz) -> dict:
    fq = K.generate(cfg, scale=scale, seed=seed)
    data = gzip.compress(fq, compresslevel=6, mtime=0) if gz else fq
    with tempfile.TemporaryDirectory(dir="/dev/shm") as td:
        td = Path(td)
        path = td / ("in.fq.gz" if gz else "in.fq")
        path.write_bytes(data)
        (td / "tmp").mkdir()
        seen = []
        for attempt in range(6):
            r = subprocess.run([str(REF / "Tree"), str(path), str(k), str(n_thread), "8"], cwd=td,
                               capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"{name}: Tree failed: {r.stderr[-400:]}")
            got = stage0_lines(r.stdout)
            if got in seen:
                break
            seen.append(got)
        else:
            raise RuntimeError(f"{name}: no two reference runs agreed on Stage 0")
    return {"workload": cfg, "scale": scale, "seed": seed, "k": k, "n_thread": n_thread, "gzip": gz,
            "file_size": len(data), "fastq_bytes": len(fq), **got}


def main() -> None:
    if not (REF / "Tree").exists():
        sys.exit("oracle/_ref/Tree missing: run `make -C oracle` in the build container")
    only = set(sys.argv[1:])
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for case in CASES:
        if only and case[0] not in only:
            continue
        print(f"{case[0]} ...", flush=True)
        data[case[0]] = run_case(*case)
        print(f"  {data[case[0]]}", flush=True)
        OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
