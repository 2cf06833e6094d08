"""Plants plausible mistakes in copies of the oracle and checks that the
oracle's pin tests fail on each one (run by hand: python tests/oracle_mutations.py).

Each entry edits one line of a copy of oracle/*.cpp (the sources in oracle/
are never touched), builds the copy into a temporary directory and runs
tests/test_oracle.py and tests/test_oracle_pins.py against it through
DGPU_ORACLE_LIB.  A mistake that no pin catches is reported as MISSED and
makes the script exit non-zero.
"""
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (file, original text, planted text, what the mistake is)
MUTATIONS = [
    ("oracle_rk.cpp", "cf1[ci(i, k, j)] += (p.ro_inv / p.delta) * cu3",
     "cf1[ci(i, k, j)] -= (p.ro_inv / p.delta) * cu3", "Coriolis sense of rotation"),
    ("oracle_rk.cpp", "cth[ci(i, k, j)] * dyf[j - 1] + cth[ci(i, k, j - 1)] * dyf[j]",
     "cth[ci(i, k, j)] * dyf[j] + cth[ci(i, k, j - 1)] * dyf[j - 1]", "buoyancy interpolation weights swapped"),
    ("oracle_rk.cpp", "for (int i = (k == 0 ? 1 : 0); i < Nxp; ++i)", "for (int i = 0; i < Nxp; ++i)",
     "buoyancy applied to the mean mode"),
    ("oracle_rk.cpp", "dyf[j - 1] * U1[ri(i, k, j - 1)]) / (2.0 * dy[j])) *",
     "dyf[j - 1] * U1[ri(i, k, j - 1)]) / (2.0 * dyf[j])) *", "GY interpolation of u1 over 2 dyf"),
    ("oracle_rk.cpp", "return (0.25 * (a * a) - 0.25 * (b * b)) / dy[j];",
     "return (0.25 * (a * a) - 0.25 * (b * b)) / dyf[j];", "u2 u2 flux over dyf"),
    ("oracle_rk.cpp", "U1[ri(i, k, j)] * U2[ri(i, k, j)] - U1[ri(i, k, j - 1)] * U2[ri(i, k, j)]) /",
     "U1[ri(i, k, j)] * U2[ri(i, k, j)] - U1[ri(i, k, j - 1)] * U2[ri(i, k, j - 1)]) /",
     "u1 u2 flux with u2 from the wrong plane"),
    ("oracle_rk.cpp", "cr1[ci(i, k, j)] -= temp4 * ik(kx[i], cp[ci(i, k, j)]);",
     "cr1[ci(i, k, j)] += temp4 * ik(kx[i], cp[ci(i, k, j)]);", "sign of the x pressure gradient"),
    ("oracle_rk.cpp", "- cu1[ci(i, k, j)] * p.dthdx[n] - cu3[ci(i, k, j)] * p.dthdz[n];",
     "- cu3[ci(i, k, j)] * p.dthdx[n] - cu1[ci(i, k, j)] * p.dthdz[n];", "background gradients paired with u3, u1"),
    ("oracle.cpp", "row(1, 0.0, -1.0, 1.0, dy[2] * c1l);", "row(1, 0.0, -1.0, 1.0, dyf[1] * c1l);",
     "lower Neumann row with dyf(1)"),
    ("oracle.cpp", "row(Nyp, -1.0, 1.0, 0.0, dy[Nyp] * c1u);", "row(Nyp, -1.0, 1.0, 0.0, dyf[Nyp] * c1u);",
     "upper Neumann row with dyf(Ny)"),
    ("oracle.cpp", "cu1[ci(0, 0, 1)] = cu1[ci(0, 0, 2)] - dy[2] * p.u_bc_ymin_c1;",
     "cu1[ci(0, 0, 1)] = cu1[ci(0, 0, 2)] + dy[2] * p.u_bc_ymin_c1;", "sign of the Neumann post BC"),
    ("oracle_rk.cpp", "(dx * (gyf[j] - 0.5 * Ly) * dl / ro) * ik(kz[k], cu1[ci(i, k, j)])",
     "(dx * (gy[j] - 0.5 * Ly) * dl / ro) * ik(kz[k], cu1[ci(i, k, j)])", "Front wind at GY for u1"),
    ("oracle_rk.cpp", "(-1.0 * dz * dl / ro) * u2m;", "(dz * dl / ro) * u2m;", "sign of the Front v term"),
]


def main():
    missed = 0
    with tempfile.TemporaryDirectory() as tmp:
        shutil.copytree(os.path.join(ROOT, "include"), os.path.join(tmp, "include"))
        for f, old, new, what in MUTATIONS:
            src = os.path.join(tmp, "oracle")
            shutil.rmtree(src, ignore_errors=True)
            shutil.copytree(os.path.join(ROOT, "oracle"), src,
                            ignore=shutil.ignore_patterns("*.so", "_ref"))
            text = open(os.path.join(src, f)).read()
            if old not in text:
                raise SystemExit(f"mutation '{what}': original text not found in {f}")
            open(os.path.join(src, f), "w").write(text.replace(old, new, 1))
            lib = os.path.join(tmp, "liboracle_mut.so")
            subprocess.run(["g++", "-O2", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-o", lib, "oracle.cpp",
                            "oracle_rk.cpp", "oracle_les.cpp", "oracle_init.cpp"], cwd=src, check=True)
            env = dict(os.environ, DGPU_ORACLE_LIB=lib)
            r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                                "tests/test_oracle.py", "tests/test_oracle_pins.py"],
                               cwd=ROOT, env=env, capture_output=True, text=True)
            failed = [line for line in r.stdout.splitlines() if line.startswith("FAILED")]
            if r.returncode == 0:
                missed += 1
                print(f"MISSED  {what}", flush=True)
            else:
                print(f"caught  {what}: {failed[0].split(' ')[1] if failed else r.stdout[-200:]}", flush=True)
    sys.exit(1 if missed else 0)


if __name__ == "__main__":
    main()
