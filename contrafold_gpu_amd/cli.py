"""Batch RNA folding command line on the GPU fold engine.

The command-line contract of the reference CLI (gpu_contrafold/cli.py:1-230):
the same positional input (a literal sequence, or a JSONL / FASTA /
one-sequence-per-line file), the same flags, the same stdout and JSONL
records, and the same exit messages.

    python -m contrafold_gpu_amd GGGGAAAACCCC               # MEA structure (gamma 6)
    python -m contrafold_gpu_amd GGGGAAAACCCC --gamma 4
    python -m contrafold_gpu_amd GGGGAAAACCCC --viterbi     # Viterbi / MAP structure
    python -m contrafold_gpu_amd GGGGAAAACCCC --sample 10   # Boltzmann samples
    python -m contrafold_gpu_amd GGGGAAAACCCC --logz        # log partition function
    python -m contrafold_gpu_amd seqs.jsonl -o out.jsonl    # JSONL / FASTA / one sequence per line

CONTRAfold-binary output formats (gpu_contrafold/reference.py:61-132):
``--posteriors CUTOFF`` (sparse 'i base j:p' lines), ``--parens`` (>id,
sequence, structure) and ``--partition`` ('Log partition coefficient' lines).

Every mode folds on the GPU in length-sorted batches of ``--chunk`` records
(the reference runs MEA and Viterbi per record on the CPU, cli.py:188-193);
``--threads`` is accepted and ignored (launch shapes come from the tuner).
"""
import argparse
import json
import os
import sys
from collections import namedtuple

import numpy as np

Record = namedtuple("Record", "id seq constrain")

RNA_LETTERS = frozenset("ACGUTN")


# ------------------------------------------------------------------ input
def _stripped(path):
    """(line number, stripped text) of every non-blank line of a file."""
    with open(path) as fh:
        for no, text in enumerate(fh):
            text = text.strip()
            if text:
                yield no, text


def sniff(path):
    """Input format of a file from its first non-blank character:
    '{' JSONL, '>' FASTA, anything else one sequence per line."""
    for _no, text in _stripped(path):
        return {"{": "jsonl", ">": "fasta"}.get(text[0], "lines")
    return "lines"


def _jsonl(path):
    """JSONL records: 'seq' required, 'id' defaults to the line number,
    'constrain' optional (a 0/1 list or string)."""
    out = []
    for no, text in _stripped(path):
        try:
            obj = json.loads(text)
        except json.JSONDecodeError as e:
            raise SystemExit(f"line {no}: invalid JSON ({e})")
        seq = obj.get("seq")
        if not seq:
            raise SystemExit(f"line {no}: missing required field 'seq'")
        out.append(Record(obj.get("id", no), str(seq).strip().upper(), obj.get("constrain")))
    return out


def _fasta(path):
    """FASTA records: a '>' line names the record (an empty name becomes the
    record's index), the following lines are its sequence."""
    names, bodies = [], []
    for _no, text in _stripped(path):
        if text.startswith(">"):
            names.append(text[1:].strip() or sum(1 for b in bodies if b))
            bodies.append([])
        elif bodies:
            bodies[-1].append(text)
        else:                           # sequence before any header
            names.append(None)
            bodies.append([text])
    return [Record(n, "".join(b).upper(), None) for n, b in zip(names, bodies) if b]


def _lines(path):
    return [Record(no, text.upper(), None) for no, text in _stripped(path)]


def load_records(src):
    """(records, from_file) for a literal sequence or a file path."""
    if os.path.isfile(src):
        reader = {"jsonl": _jsonl, "fasta": _fasta, "lines": _lines}[sniff(src)]
        return reader(src), True
    seq = src.strip().upper()
    if seq and RNA_LETTERS.issuperset(seq):
        return [Record(0, seq, None)], False
    raise SystemExit(f"'{src}' is not an existing file nor a valid RNA sequence (ACGUTN)")


def build_mask(seq, constrain, where=""):
    """Forced-unpaired mask (int8, 1 = forced) from a 0/1 list or a '0'/'1'
    string of the sequence's length; None when there is no constraint."""
    if constrain is None or len(constrain) == 0:
        return None
    if isinstance(constrain, str):
        vals = np.frombuffer(constrain.encode("ascii", "replace"), np.uint8) == ord("1")
        vals = vals.astype(np.int64)
    else:
        vals = list(constrain)
    if len(vals) != len(seq):
        raise SystemExit(f"{where}constrain length {len(vals)} != seq length {len(seq)}")
    arr = np.asarray(vals, dtype=object)
    bad = [k for k, v in enumerate(arr) if v not in (0, 1)]
    if bad:
        raise SystemExit(f"{where}constrain[{bad[0]}] = {arr[bad[0]]!r}, expected 0 or 1")
    return arr.astype(np.int8)


# ------------------------------------------------------------------ folding
def fold(records, P, mode="mea", gamma=6.0, sample_n=0, with_logz=False, chunk=4096, seed=0):
    """Fold records in length-sorted batches of ``chunk``; results in input
    order: {'structure': [...], 'samples': [...], 'logZ': [...]}."""
    import contrafold_gpu_amd as cf
    eng = cf.default_engine()
    n = len(records)
    masks = [build_mask(r.seq, r.constrain, where=f"id={r.id!r}: ") for r in records]
    res = {"structure": [None] * n, "samples": [None] * n, "logZ": [None] * n}
    flags = {"mea": cf.CF_MEA, "mfe": cf.CF_MFE, "sample": cf.CF_SAMPLE, "logz": cf.CF_LOGZ}[mode]
    by_length = sorted(range(n), key=lambda k: len(records[k].seq))
    step = max(1, int(chunk))
    with eng.lock:
        eng.set_params(P)
        for start in range(0, n, step):
            part = by_length[start:start + step]
            with eng.batch([records[k].seq for k in part], [masks[k] for k in part]) as b:
                b.run(flags, gammas=[gamma] if mode == "mea" else (), n_samples=sample_n, seed=seed + start)
                if mode == "mea":
                    got = b.mea(0)
                elif mode == "mfe":
                    got = b.mfe()[0]
                elif mode == "sample":
                    got = b.samples()
                else:
                    got = None
                key = "samples" if mode == "sample" else "structure"
                if got is not None:
                    for k, v in zip(part, got):
                        res[key][k] = v
                if mode == "logz" or with_logz:
                    for k, z in zip(part, b.logz()):
                        res["logZ"][k] = float(z)
    return res


# ------------------------------------------------------------------ output
def write_posteriors(records, P, cutoff, out):
    """CONTRAfold --posteriors text: one line per position i (1-based),
    'i base j:p ...' listing every partner j (either side) with p >= cutoff;
    records are separated by a '>id' line.  The reference parses exactly this
    (gpu_contrafold/reference.py:103-132)."""
    import contrafold_gpu_amd as cf
    masks = [build_mask(r.seq, r.constrain, where=f"id={r.id!r}: ") for r in records]
    mats = cf.bpp_batch([r.seq for r in records], P, masks)
    for rec, M in zip(records, mats):
        full = M + M.T
        if len(records) > 1:
            out.write(f">{rec.id}\n")
        for i, base in enumerate(rec.seq):
            toks = [f"{j + 1}:{full[i, j]:.6g}" for j in np.nonzero(full[i] >= cutoff)[0] if j != i]
            out.write(" ".join([str(i + 1), base] + toks) + "\n")


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


def _json_record(mode, rec_id, res, k, with_logz):
    row = {"id": rec_id}
    if mode in ("mea", "mfe"):
        row["structure"] = res["structure"][k]
    elif mode == "sample":
        if with_logz:
            row["logZ"] = res["logZ"][k]
        row["samples"] = res["samples"][k]
    else:
        row["logZ"] = res["logZ"][k]
    return row


def _print_literal(mode, res, with_logz):
    """Human-readable stdout for one literal sequence without -o."""
    if mode in ("mea", "mfe"):
        print(res["structure"][0])
    elif mode == "sample":
        if with_logz:
            print(f"# logZ = {res['logZ'][0]:.6f}")
        print("\n".join(res["samples"][0]))
    else:
        print(f"{res['logZ'][0]:.6f}")


class _Sink:
    """The -o file, or stdout; closes only what it opened."""

    def __init__(self, path):
        self.path = path
        self.fh = open(path, "w") if path else sys.stdout

    def __enter__(self):
        return self.fh

    def __exit__(self, *exc):
        if self.path:
            self.fh.close()


def _parser():
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
    return ap


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    # an optional leading 'fold' sub-command is accepted and dropped
    args = _parser().parse_args(argv[1:] if argv[:1] == ["fold"] else argv)
    records, from_file = load_records(args.input)
    if not records:
        raise SystemExit("no sequences in input")
    import contrafold_gpu_amd as cf
    P = cf.load()
    if args.posteriors is not None:
        with _Sink(args.output) as out:
            write_posteriors(records, P, args.posteriors, out)
        return
    if args.partition:
        res = fold(records, P, mode="logz", chunk=args.chunk)
        with _Sink(args.output) as out:
            for rec, z in zip(records, res["logZ"]):
                out.write(f'Log partition coefficient for "{rec.id}": {z:.6g}\n')
        return
    if args.sample:
        mode = "sample"
    elif args.logz:
        mode = "logz"
    else:
        mode = "mfe" if args.viterbi else "mea"
    with_logz = args.logz and mode == "sample"
    res = fold(records, P, mode=mode, gamma=args.gamma, sample_n=args.sample, with_logz=with_logz,
               chunk=args.chunk, seed=args.seed)
    if not from_file and not args.output:
        _print_literal(mode, res, with_logz)
        return
    with _Sink(args.output) as out:
        if args.parens and mode in ("mea", "mfe"):
            for rec, db in zip(records, res["structure"]):
                out.write(f">{rec.id}\n{rec.seq}\n{db}\n")
        else:
            for k, rec in enumerate(records):
                out.write(json.dumps(_json_record(mode, rec.id, res, k, args.logz)) + "\n")
    if args.output and not args.parens:
        print(f"[contrafold-amd] wrote {len(records)} records -> {args.output}", file=sys.stderr)


if __name__ == "__main__":
    main()
