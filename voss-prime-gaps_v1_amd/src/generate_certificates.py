#!/usr/bin/env python3
"""Certificate stage of the MI355X build (drop-in for the reference's
scripts/generate_certificates.py, run by run_voss.py as step [4/8]).

Input: results/large_gaps.csv (position = p_after, gap, merit), relative to
the parent of this file's directory, as in the reference.  Output format
(scripts/generate_certificates.py:105-108, :196-233, README.md:277-278):

  results/certificates.csv   p_before,p_after,gap,merit,sha256,verified: one
                             row per input row in file order; merit printed
                             with 6 decimals; sha256 = SHA-256 of the text
                             "VOSS|p_before|p_after|gap|merit" (the same
                             6-decimal merit); verified YES or NO
  results/cert_top20/        text certificates of the 20 largest merits
                             (ties keep file order), named
                             cert_NN_merit<merit, 2 decimals>_p<p_after>.txt

Where verification flags come from: results/verification_details.csv (the
verifier's output) for every (p_after, gap) it lists; the rest are checked
by this build's verify_cpp (GPU Miller-Rabin on both ends and on the
interior numbers) in one call.  The reference checks those with numba
functions instead; numba plays no part here.
"""
import csv
import hashlib
import math
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.abspath(os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
RESULTS = os.path.join(REPO, "results")
GAPS_FILE = os.path.join(RESULTS, "large_gaps.csv")
DETAILS_FILE = os.path.join(RESULTS, "verification_details.csv")
CSV_OUT = os.path.join(RESULTS, "certificates.csv")
TEXT_DIR = os.path.join(RESULTS, "cert_top20")
TOP = 20

RULE, THIN = "=" * 66, "-" * 66
# the text certificate, line for line the reference's layout
TEXT_LAYOUT = "\n".join([
    RULE,
    "           VOSS NUMERICAL CERTIFICATE OF PRIME GAP",
    RULE,
    "Prime Before (p)        : {pb:,}",
    "Prime After  (p+g)      : {pa:,}",
    "Gap Size (g)            : {g:,}",
    "Merit (g / ln p)        : {merit:.6f}",
    "ln(p)                   : {lnp:.6f}",
    "Max Factor Limit (sqrtp+g) : {root:,}",
    THIN,
    "VERIFICATION METHOD:",
    "  1. p_before and p_after are proven prime via",
    "     Deterministic Miller-Rabin (bases 2,3,5,7,11,13,17,19,23,29,31,37).",
    "  2. All {inner} integers in (p, p+g) are proven composite",
    "     via Interval Sieve with base primes up to √(p+g).",
    THIN,
    "SHA-256: {sha}",
    "Generated: {stamp}",
    RULE,
    "",
])


def record_digest(p_before, p_after, gap, merit):
    """SHA-256 (hex) of the certified record VOSS|p_before|p_after|gap|merit."""
    record = "|".join(("VOSS", str(p_before), str(p_after), str(gap), format(merit, ".6f")))
    return hashlib.sha256(record.encode()).hexdigest()


def read_gaps(path):
    """[(p_after, gap, merit)] in file order, by the header's column names."""
    with open(path, newline="") as f:
        table = csv.reader(f)
        head = next(table, [])
        col = {name: k for k, name in enumerate(head)}
        return [(int(r[col["position"]]), int(r[col["gap"]]), float(r[col["merit"]]))
                for r in table if r]


def read_flags(path):
    """{(p_after, gap): verified} from a verifier details file."""
    flags = {}
    with open(path, newline="") as f:
        table = csv.reader(f)
        head = next(table, [])
        col = {name: k for k, name in enumerate(head)}
        for r in table:
            if r:
                flags[(int(r[col["position"]]), int(r[col["gap"]]))] = r[col["verified"]] == "1"
    return flags


def verifier_program():
    """bin/verify_cpp of this build, under $VOSS_AMD_HOME or this tree's bin/."""
    home = os.environ.get("VOSS_AMD_HOME")
    places = []
    if home:
        places += [os.path.join(home, "voss-prime-gaps_v1_amd", "bin", "verify_cpp"),
                   os.path.join(home, "bin", "verify_cpp")]
    places.append(os.path.join(REPO, "bin", "verify_cpp"))
    return next((p for p in places if os.path.isfile(p) and os.access(p, os.X_OK)), None)


def check_on_gpu(gaps):
    """Verification flags of (p_after, gap, merit) rows from one verify_cpp
    run over a temporary list (its own empty cache)."""
    exe = verifier_program()
    if exe is None:
        sys.exit("gaps without verification details, and verify_cpp is not built "
                 "(build voss-prime-gaps_v1_amd, set VOSS_AMD_HOME)")
    with tempfile.TemporaryDirectory() as scratch:
        listed = os.path.join(scratch, "gaps.csv")
        details = os.path.join(scratch, "details.csv")
        with open(listed, "w") as f:
            f.write("position,gap,merit\n")
            f.writelines(f"{pa},{g},{m:.6f}\n" for pa, g, m in gaps)
        subprocess.run([exe, listed, details, os.path.join(scratch, "cache.csv")],
                       stdout=subprocess.DEVNULL)
        return read_flags(details)


def certificate_text(p_after, gap, merit, sha):
    pb = p_after - gap
    return TEXT_LAYOUT.format(pb=pb, pa=p_after, g=gap, merit=merit, lnp=math.log(pb),
                              root=math.isqrt(p_after), inner=gap - 1, sha=sha,
                              stamp=time.strftime("%Y-%m-%d %H:%M:%S"))


def main():
    if not os.path.exists(GAPS_FILE):
        print(f"{GAPS_FILE} not found")
        return 1
    gaps = read_gaps(GAPS_FILE)
    print(f"{len(gaps):,} gaps in {GAPS_FILE}")
    if not gaps:
        return 0
    started = time.time()
    flags = read_flags(DETAILS_FILE) if os.path.exists(DETAILS_FILE) else {}
    if flags:
        print(f"verification flags of {len(flags):,} gaps from {DETAILS_FILE}")
    unknown = [row for row in gaps if (row[0], row[1]) not in flags]
    if unknown:
        print(f"checking {len(unknown):,} gaps on the GPU")
        flags.update(check_on_gpu(unknown))

    os.makedirs(RESULTS, exist_ok=True)
    yes = 0
    with open(CSV_OUT, "w", newline="") as f:
        out = csv.writer(f)
        out.writerow(["p_before", "p_after", "gap", "merit", "sha256", "verified"])
        for p_after, gap, merit in gaps:
            ok = flags.get((p_after, gap), False)
            yes += ok
            out.writerow([p_after - gap, p_after, gap, format(merit, ".6f"),
                          record_digest(p_after - gap, p_after, gap, merit),
                          "YES" if ok else "NO"])
    print(f"{CSV_OUT}: {len(gaps):,} certificates, {yes:,} verified "
          f"({time.time() - started:.2f} s)")

    os.makedirs(TEXT_DIR, exist_ok=True)
    best = sorted(gaps, key=lambda row: row[2], reverse=True)[:TOP]  # stable: ties in file order
    for rank, (p_after, gap, merit) in enumerate(best, 1):
        name = f"cert_{rank:02d}_merit{merit:.2f}_p{p_after}.txt"
        sha = record_digest(p_after - gap, p_after, gap, merit)
        with open(os.path.join(TEXT_DIR, name), "w", encoding="utf-8") as f:
            f.write(certificate_text(p_after, gap, merit, sha))
        print(f"  {name}")
    print(f"{TEXT_DIR}: {len(best)} text certificates (largest merits)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
