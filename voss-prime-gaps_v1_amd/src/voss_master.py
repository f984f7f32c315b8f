N_VALUE = 100000000000      # 10^11 -- run_voss.py rewrites the first N_VALUE
SEG_NUM = 30000000000       # 3x10^10 -- and the first SEG_NUM assignment
# ============================================================
# VOSS master, MI355X drop-in.  Replaces the reference's src/voss_master.py:
# instead of patching constants into a CUDA source and compiling it with
# nvcc (src/voss_master.py:64-85, :138-143), it runs the prebuilt gfx950
# program voss_v7_live with N and SEG_NUM as arguments, cwd = results/,
# streams its output and checks "Total primes:" against the OEIS table in
# data/verification_db.json (:150-179).  The two assignments above must stay
# the first lines that match run_voss.py's regexes (run_voss.py:118-121).
# ============================================================
import json
import math
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, '..'))
RESULTS = os.path.join(ROOT, 'results')
DB_FILE = os.path.join(ROOT, 'data', 'verification_db.json')


def find_binary(name):
    """The program built by voss-prime-gaps_v1_amd/build.py: from
    $VOSS_AMD_HOME, from this tree's bin/, or beside this package."""
    cands = []
    home = os.environ.get('VOSS_AMD_HOME')
    if home:
        cands += [os.path.join(home, 'voss-prime-gaps_v1_amd', 'bin', name),
                  os.path.join(home, 'bin', name)]
    cands += [os.path.join(ROOT, 'bin', name)]
    for c in cands:
        if os.path.isfile(c) and os.access(c, os.X_OK):
            return c
    return None


def load_db():
    if not os.path.exists(DB_FILE):
        print('No ' + DB_FILE + ' (no download is attempted); OEIS check skipped')
        return {}
    with open(DB_FILE) as f:
        db = json.load(f)
    print('DB OK: ' + DB_FILE)
    return db.get('pi_of_10_pow_n', {})


def main():
    print('=' * 60)
    print('VOSS Master (MI355X)')
    print('=' * 60)
    print(f'ROOT  = {ROOT}')
    binary = find_binary('voss_v7_live')
    if binary is None:
        print('ERROR: voss_v7_live not found (build voss-prime-gaps_v1_amd and set VOSS_AMD_HOME)')
        sys.exit(1)
    print(f'Binary = {binary}')
    print()
    pi_db = load_db()
    n_log10 = round(math.log10(N_VALUE))
    is_pow10 = abs(10**n_log10 - N_VALUE) / N_VALUE < 1e-9
    expected_pi = pi_db.get(str(n_log10)) if is_pow10 else None
    if expected_pi:
        print(f'pi(10^{n_log10}) from OEIS = {expected_pi:,}')
    print()
    print(f'N={N_VALUE:,}  SEG_NUM={SEG_NUM:,}  NUM_SEG={math.ceil(N_VALUE / SEG_NUM)}')
    print('=' * 60)
    os.makedirs(RESULTS, exist_ok=True)
    print(f'Running in: {RESULTS}\n')
    proc = subprocess.Popen([binary, str(N_VALUE), str(SEG_NUM)], cwd=RESULTS,
                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, bufsize=0)
    captured = []
    try:
        while True:
            ch = proc.stdout.read(1)
            if not ch:
                break
            t = ch.decode('utf-8', errors='replace')
            captured.append(t)
            sys.stdout.write(t)
            sys.stdout.flush()
    except KeyboardInterrupt:
        proc.terminate()
    rc = proc.wait()
    full = ''.join(captured)
    print('\n' + '=' * 60)
    print('=== AUTO-VERIFICATION vs OEIS A006880 ===')
    print('=' * 60)
    m = re.search(r'Total primes:\s*(\d+)', full)
    actual = int(m.group(1)) if m else None
    if actual and expected_pi:
        print(f'OEIS pi(10^{n_log10}) = {expected_pi:>25,}')
        print(f'VOSS pi(10^{n_log10}) = {actual:>25,}')
        d = actual - expected_pi
        print(f'Diff = {d:+,}')
        print('PASS' if d == 0 else f'FAIL ({d:,})')
    else:
        print('Verification skipped (not a power of 10, or parse failed).')
    if rc != 0:
        print(f'voss_v7_live exited with code {rc}')
        sys.exit(rc)


if __name__ == '__main__':
    main()
