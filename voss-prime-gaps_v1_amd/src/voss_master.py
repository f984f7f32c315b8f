N_VALUE = 100000000000      # 10^11 -- run_voss.py rewrites the first N_VALUE
SEG_NUM = 30000000000       # 3x10^10 -- and the first SEG_NUM assignment
"""VOSS driver for the MI355X build (drop-in for the reference's
src/voss_master.py).

run_voss.py (run_voss.py:111-123, :198) rewrites the first `N_VALUE = <n>` and
`SEG_NUM = <n>` assignments of this file, runs it with the repository root as
working directory, streams its output and stops the pipeline on a non-zero
exit code; afterwards it reads results/large_gaps.csv.  Those two
assignments therefore stay the first lines of the file.

What this driver does: it runs the prebuilt gfx950 program voss_v7_live with
N and SEG_NUM as arguments (nothing is patched or compiled), with results/ as
its working directory, where the program writes its CSV files; it passes the
program's output through as it arrives (the progress line rewrites itself
with carriage returns); it compares the program's "Total primes: <pi>" with
pi(10^k) from data/verification_db.json when N is a power of ten (no
download is attempted); and it exits with the program's exit code, so a
failed run stops run_voss.py instead of leaving stale result files behind.
"""
import json
import os
import re
import subprocess
import sys

REPO = os.path.abspath(os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
OUT_DIR = os.path.join(REPO, "results")
PI_TABLE = os.path.join(REPO, "data", "verification_db.json")
PROGRAM = "voss_v7_live"


def program_path():
    """voss_v7_live as built by voss-prime-gaps_v1_amd/build.py: under
    $VOSS_AMD_HOME (the MI355X package or its bin/), else this tree's bin/."""
    places = []
    home = os.environ.get("VOSS_AMD_HOME")
    if home:
        places.append(os.path.join(home, "voss-prime-gaps_v1_amd", "bin", PROGRAM))
        places.append(os.path.join(home, "bin", PROGRAM))
    places.append(os.path.join(REPO, "bin", PROGRAM))
    return next((p for p in places if os.path.isfile(p) and os.access(p, os.X_OK)), None)


def power_of_ten(n):
    """k when n == 10**k, else None."""
    k, v = 0, 1
    while v < n:
        v *= 10
        k += 1
    return k if v == n else None


def published_pi(n):
    """pi(n) from the local OEIS A006880 table when n is a power of ten."""
    k = power_of_ten(n)
    if k is None or not os.path.isfile(PI_TABLE):
        return None
    with open(PI_TABLE) as f:
        table = json.load(f).get("pi_of_10_pow_n", {})
    v = table.get(str(k))
    return int(v) if v is not None else None


def run_streamed(argv, cwd):
    """Runs argv in cwd, copying its merged stdout/stderr to ours as it
    comes; returns (exit code, everything it printed)."""
    seen = []
    child = subprocess.Popen(argv, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    try:
        while True:
            chunk = os.read(child.stdout.fileno(), 4096)
            if not chunk:
                break
            seen.append(chunk)
            sys.stdout.buffer.write(chunk)
            sys.stdout.flush()
    except KeyboardInterrupt:
        child.terminate()
    code = child.wait()
    return code, b"".join(seen).decode("utf-8", errors="replace")


def main():
    n, seg = int(N_VALUE), int(SEG_NUM)
    exe = program_path()
    bar = "-" * 60
    print(bar)
    print(f"VOSS driver (MI355X build)   N = {n:,}   SEG_NUM = {seg:,}   "
          f"logical segments = {-(-n // seg)}")
    print(bar)
    if exe is None:
        print(f"{PROGRAM} is not built: run voss-prime-gaps_v1_amd/build.py and point "
              "VOSS_AMD_HOME at it")
        return 1
    want = published_pi(n)
    os.makedirs(OUT_DIR, exist_ok=True)
    print(f"program  {exe}\noutputs  {OUT_DIR}\n", flush=True)
    code, text = run_streamed([exe, str(n), str(seg)], OUT_DIR)
    hit = re.search(r"Total primes:\s*(\d+)", text)
    got = int(hit.group(1)) if hit else None
    print("\n" + bar)
    if got is not None and want is not None:
        verdict = "PASS" if got == want else f"FAIL (off by {got - want:+,})"
        print(f"pi({n:,}) = {got:,}; OEIS A006880 gives {want:,}: {verdict}")
    elif got is not None:
        print(f"pi({n:,}) = {got:,} (no published value to compare with)")
    else:
        print("the program printed no 'Total primes:' line")
    if code != 0:
        print(f"{PROGRAM} exited with code {code}")
    print(bar)
    return code


if __name__ == "__main__":
    sys.exit(main())
