"""Compare two vis.out files for numerical agreement.

    python -m gcorr_amd.verify --config CONF [--maxepsilon E] [--autos] REF.out NEW.out

The command-line form of fxkernel/scripts/verifyfxkernel.py:33-77: the
.conf gives the station count, and every cross baseline (plus the
autocorrelation rows with --autos) and polarisation product is checked.
The measure is visio.compare_text's: per row and product, the largest
complex difference over channels divided by the median reference
amplitude.  Prints one FAIL line per series above --maxepsilon, or PASS with
the largest absolute difference; exits 0 on PASS, 1 on FAIL, 2 on unusable
input.
"""
import argparse
import os
import sys

from . import visio


def main(argv=None):
    ap = argparse.ArgumentParser(description="Compare two vis.out files for numerical correctness")
    ap.add_argument("--config", required=True, help=".conf with NANT (and NPOL)")
    ap.add_argument("--maxepsilon", type=float, default=1e-7,
                    help="largest allowed difference as a fraction of the median visibility amplitude")
    ap.add_argument("--autos", action="store_true", help="the files carry one autocorrelation row per station")
    ap.add_argument("visfiles", nargs=2, help="reference and new vis.out")
    args = ap.parse_args(argv)
    for f in [args.config] + args.visfiles:
        if not os.path.exists(f):
            print("%s doesn't exist" % f, file=sys.stderr)
            return 2
    nant = npol = None
    with open(args.config) as fh:
        for line in fh:
            w = line.split()
            if len(w) >= 2 and w[0].upper() == "NANT":
                nant = int(w[1])
            elif len(w) >= 2 and w[0].upper() == "NPOL":
                npol = int(w[1])
    if nant is None:
        print("no NANT in %s" % args.config, file=sys.stderr)
        return 2
    nprod = (npol or 2) ** 2
    nrows = nant * (nant - 1) // 2 + (nant if args.autos else 0)
    try:
        r = visio.compare_text(args.visfiles[0], args.visfiles[1], nrows, args.maxepsilon, nprod)
    except ValueError as e:
        print(str(e), file=sys.stderr)
        return 2
    for row, prod, frac in r["fail"]:
        print("FAIL: Baseline {0} pol product {1} has a max fractional difference of {2}".format(row, prod, frac))
    if r["pass"]:
        print("PASS: maximum absolute difference on any baseline was", r["max_abs_diff"])
        return 0
    return 1


if __name__ == "__main__":
    sys.exit(main())
