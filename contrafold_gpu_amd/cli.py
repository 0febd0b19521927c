"""Batch RNA folding command line on the GPU fold engine.

Same inputs, modes and outputs as the reference CLI (gpu_contrafold/cli.py:1-230):

    python -m contrafold_gpu_amd GGGGAAAACCCC               # MEA structure (gamma 6)
    python -m contrafold_gpu_amd GGGGAAAACCCC --gamma 4
    python -m contrafold_gpu_amd GGGGAAAACCCC --viterbi     # Viterbi / MAP structure
    python -m contrafold_gpu_amd GGGGAAAACCCC --sample 10   # Boltzmann samples
    python -m contrafold_gpu_amd GGGGAAAACCCC --logz        # log partition function
    python -m contrafold_gpu_amd seqs.jsonl -o out.jsonl    # JSONL / FASTA / one sequence per line

Differences from the reference: MEA and Viterbi modes fold on the GPU in
length-sorted batches (the reference folds them one record at a time on the
CPU, cli.py:188-193); ``--threads`` is accepted and ignored (launch shapes come
from the tuner).
"""
import argparse
import json
import os
import sys

import numpy as np

_SEQ_CHARS = set("ACGUTN")


def _parse_jsonl(path):
    recs = []
    with open(path) as fh:
        for ln, line in enumerate(fh):
            line = line.strip()
            if not line:
                continue
            try:
                obj = json.loads(line)
            except json.JSONDecodeError as e:
                raise SystemExit(f"line {ln}: invalid JSON ({e})")
            if "seq" not in obj or not obj["seq"]:
                raise SystemExit(f"line {ln}: missing required field 'seq'")
            recs.append((obj.get("id", ln), str(obj["seq"]).strip().upper(), obj.get("constrain")))
    return recs


def _parse_fasta(path):
    recs, rid, cur = [], None, []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith(">"):
                if cur:
                    recs.append((rid, "".join(cur).upper(), None))
                    cur = []
                rid = line[1:].strip() or len(recs)
            elif line.strip():
                cur.append(line.strip())
    if cur:
        recs.append((rid, "".join(cur).upper(), None))
    return recs


def load_records(src):
    """(records, from_file); records are (id, seq, constrain)."""
    if os.path.isfile(src):
        head = ""
        with open(src) as fh:
            for line in fh:
                if line.strip():
                    head = line.strip()[0]
                    break
        if head == "{":
            return _parse_jsonl(src), True
        if head == ">":
            return _parse_fasta(src), True
        with open(src) as fh:
            return [(i, ln.strip().upper(), None) for i, ln in enumerate(fh) if ln.strip()], True
    seq = src.strip().upper()
    if seq and all(c in _SEQ_CHARS for c in seq):
        return [(0, seq, None)], False
    raise SystemExit(f"'{src}' is not an existing file nor a valid RNA sequence (ACGUTN)")


def build_mask(seq, constrain, where=""):
    """0/1 mask (1 = forced unpaired) from a list or a '0'/'1' string, or None."""
    if not constrain:
        return None
    L = len(seq)
    bits = [1 if c == "1" else 0 for c in constrain] if isinstance(constrain, str) else list(constrain)
    if len(bits) != L:
        raise SystemExit(f"{where}constrain length {len(bits)} != seq length {L}")
    m = np.zeros(L, np.int8)
    for i, b in enumerate(bits):
        if b not in (0, 1):
            raise SystemExit(f"{where}constrain[{i}] = {b!r}, expected 0 or 1")
        m[i] = b
    return m


def fold(records, P, mode="mea", gamma=6.0, sample_n=0, with_logz=False, chunk=4096, seed=0):
    """Fold records in length-sorted chunks; results mapped back to input order."""
    import contrafold_gpu_amd as cf
    eng = cf.default_engine()
    eng.set_params(P)
    n = len(records)
    seqs = [r[1] for r in records]
    masks = [build_mask(r[1], r[2], where=f"id={r[0]!r}: ") for r in records]
    order = sorted(range(n), key=lambda k: len(seqs[k]))
    out = {"structure": [None] * n, "samples": [None] * n, "logZ": [None] * n}
    for c0 in range(0, n, chunk):
        idx = order[c0:c0 + chunk]
        sub = [seqs[k] for k in idx]
        fl = [masks[k] for k in idx]
        with eng.batch(sub, fl) as b:
            if mode == "mea":
                b.run(cf.CF_MEA, gammas=[gamma])
                res = b.mea(0)
                for k, r in zip(idx, res):
                    out["structure"][k] = r
            elif mode == "mfe":
                b.run(cf.CF_MFE)
                res, _ = b.mfe()
                for k, r in zip(idx, res):
                    out["structure"][k] = r
            elif mode == "sample":
                b.run(cf.CF_SAMPLE, n_samples=sample_n, seed=seed + c0)
                for k, r in zip(idx, b.samples()):
                    out["samples"][k] = r
            if mode == "logz" or with_logz:
                b.run(cf.CF_LOGZ)
                for k, z in zip(idx, b.logz()):
                    out["logZ"][k] = float(z)
    return out


def write_posteriors(records, P, cutoff, out):
    """CONTRAfold --posteriors text: one line per position i (1-based),
    'i base j:p ...' listing every partner j (either side) with p >= cutoff;
    records are separated by a '>id' line.  The reference parses exactly this
    (gpu_contrafold/reference.py:103-132)."""
    import contrafold_gpu_amd as cf
    masks = [build_mask(r[1], r[2], where=f"id={r[0]!r}: ") for r in records]
    mats = cf.bpp_batch([r[1] for r in records], P, masks)
    for (rid, seq, _c), M in zip(records, mats):
        full = M + M.T
        if len(records) > 1:
            out.write(f">{rid}\n")
        for i in range(len(seq)):
            js = np.nonzero(full[i] >= cutoff)[0]
            toks = " ".join(f"{j + 1}:{full[i, j]:.6g}" for j in js if j != i)
            out.write(f"{i + 1} {seq[i]}" + (" " + toks if toks else "") + "\n")


def parse_posteriors(text, L):
    """Dense L x L upper-triangular matrix from CONTRAfold --posteriors text."""
    mat = np.zeros((L, L))
    for line in text.splitlines():
        parts = line.split()
        if len(parts) < 2 or line.startswith(">"):
            continue
        i = int(parts[0]) - 1
        for tok in parts[2:]:
            j_s, p_s = tok.split(":")
            j = int(j_s) - 1
            a, b = (i, j) if j > i else (j, i)
            mat[a, b] = float(p_s)
    return mat


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if argv and argv[0] == "fold":
        argv = argv[1:]
    ap = argparse.ArgumentParser(prog="contrafold-amd", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("input", help="RNA sequence, or path to a JSONL/FASTA/sequence-per-line file")
    ap.add_argument("-o", "--output", default=None, help="output JSONL (default: stdout)")
    ap.add_argument("--sample", type=int, default=0, metavar="N", help="draw N Boltzmann samples per sequence")
    ap.add_argument("--viterbi", "--mfe", action="store_true", dest="viterbi", help="Viterbi (MAP) structure")
    ap.add_argument("--gamma", type=float, default=6.0, metavar="G", help="MEA trade-off (default 6)")
    ap.add_argument("--logz", action="store_true", help="emit logZ instead of a structure")
    ap.add_argument("--chunk", type=int, default=4096, help="sequences per GPU batch")
    ap.add_argument("--threads", type=int, default=128, help="accepted for compatibility; ignored")
    ap.add_argument("--seed", type=int, default=0, help="RNG seed for sampling")
    ap.add_argument("--posteriors", type=float, default=None, metavar="CUTOFF",
                    help="write base-pair posteriors >= CUTOFF in CONTRAfold's sparse format "
                         "('i base j:p ...'), as `contrafold predict --posteriors`")
    ap.add_argument("--parens", action="store_true",
                    help="write structures as CONTRAfold --parens records (>id / sequence / structure)")
    ap.add_argument("--partition", action="store_true",
                    help="print CONTRAfold --partition lines (Log partition coefficient for \"id\": X)")
    args = ap.parse_args(argv)
    records, from_file = load_records(args.input)
    if not records:
        raise SystemExit("no sequences in input")
    import contrafold_gpu_amd as cf
    P = cf.load()
    if args.posteriors is not None or args.partition:
        out = open(args.output, "w") if args.output else sys.stdout
        try:
            if args.posteriors is not None:
                write_posteriors(records, P, args.posteriors, out)
            else:
                res = fold(records, P, mode="logz", chunk=args.chunk)
                for k, (rid, _s, _c) in enumerate(records):
                    out.write(f'Log partition coefficient for "{rid}": {res["logZ"][k]:.6g}\n')
        finally:
            if args.output:
                out.close()
        return
    mode = "sample" if args.sample else ("logz" if args.logz else ("mfe" if args.viterbi else "mea"))
    res = fold(records, P, mode=mode, gamma=args.gamma, sample_n=args.sample,
               with_logz=args.logz and mode == "sample", chunk=args.chunk, seed=args.seed)
    if not from_file and not args.output:
        if mode in ("mfe", "mea"):
            print(res["structure"][0])
        elif mode == "sample":
            if args.logz:
                print(f"# logZ = {res['logZ'][0]:.6f}")
            for db in res["samples"][0]:
                print(db)
        else:
            print(f"{res['logZ'][0]:.6f}")
        return
    out = open(args.output, "w") if args.output else sys.stdout
    if args.parens and mode in ("mfe", "mea"):
        try:
            for k, (rid, seq, _c) in enumerate(records):
                out.write(f">{rid}\n{seq}\n{res['structure'][k]}\n")
        finally:
            if args.output:
                out.close()
        return
    try:
        for k, (rid, _s, _c) in enumerate(records):
            rec = {"id": rid}
            if mode in ("mfe", "mea"):
                rec["structure"] = res["structure"][k]
            elif mode == "sample":
                if args.logz:
                    rec["logZ"] = res["logZ"][k]
                rec["samples"] = res["samples"][k]
            else:
                rec["logZ"] = res["logZ"][k]
            out.write(json.dumps(rec) + "\n")
    finally:
        if args.output:
            out.close()
    if args.output:
        print(f"[contrafold-amd] wrote {len(records)} records -> {args.output}", file=sys.stderr)


if __name__ == "__main__":
    main()
