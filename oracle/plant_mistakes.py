"""Check the oracle's tests: plant a typical mistake in a copy of the oracle,
build the copy, and run the CPU pin tests against it; every mistake must be
caught by at least one test. CPU only; copies and builds go to a temporary
directory.

    python oracle/plant_mistakes.py [name ...]
"""
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
src = open(f"{ROOT}/oracle/trace_oracle.cpp").read()
MUT = {
  "D without 1/pi": ("return F(0.31830988618) / (ax * ay * D_helper * D_helper);", "return T(1) / (ax * ay * D_helper * D_helper);"),
  "D alpha not squared": ("T D_helper = (H.x * H.x) / (ax * ax) + (H.y * H.y) / (ay * ay) + H.z * H.z;", "T D_helper = (H.x * H.x) / (ax) + (H.y * H.y) / (ay) + H.z * H.z;"),
  "G2 uncorrelated": ("T G2 = T(1) / (sh.one_plus_lambda_wo + ggx_lambda(wi, sh.ax, sh.ay));", "T G2 = T(1) / (sh.one_plus_lambda_wo * (T(1) + ggx_lambda(wi, sh.ax, sh.ay)));"),
  "TS factor 2": ("T ts = D * G2 / (F(4.0) * wi.z * wo.z);", "T ts = D * G2 / (F(2.0) * wi.z * wo.z);"),
  "F82 tint sign": ("V3<T> fa = s_sub(T(1), sh.specular_color) * k;", "V3<T> fa = sub_s(sh.specular_color, T(1)) * k;"),
  "E_dielectric draw order": ("T rand1 = T(rng.next());\n            T two_pi_rand = F(2.0) * F(3.14159265359) * T(rng.next());", "T two_pi_rand = F(2.0) * F(3.14159265359) * T(rng.next());\n            T rand1 = T(rng.next());"),
  "E_dielectric G1 only": ("sum_W += Fr * opl / (opl + ggx_lambda(wi, ax, ay));", "sum_W += Fr;"),
  "E_dielectric no wi.z test": ("if (wi.z <= T(0)) continue;\n            T Fr", "T Fr"),
  "sky bilinear weights swapped": ("V3<T> top = lerp3(px(xa, ya), px(xb, ya), fx);", "V3<T> top = lerp3(px(xa, ya), px(xb, ya), T(1) - fx);"),
  "sky half-texel offset dropped": ("T x = u * T(W) - F(0.5);", "T x = u * T(W);"),
  "srgb linear": ("double l = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);", "double l = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.2);"),
  "last light sample with MIS weight": ("T weight = bounce < max_bounces ? power_heuristic(light.pdf, pdf) : T(1);", "T weight = power_heuristic(light.pdf, pdf);"),
  "emission without MIS": ("emission_mis = power_heuristic(previous_bsdf_pdf, light_pdf);", "emission_mis = T(1);"),
  "roulette without division": ("                throughput = throughput / p_survive;", ""),
  "jitter y first": ("T r0 = T(rng.next());\n        T r1 = T(rng.next());", "T r1 = T(rng.next());\n        T r0 = T(rng.next());"),
  "tangent not orthogonalised": ("if (length_squared > F(1e-12)) return projected * rsqrt(length_squared);", "if (length_squared > F(1e-12)) return tangent * rsqrt(dot(tangent, tangent));"),
  "lambda missing 1/2": ("return M::fma(F(0.5), M::sqrt(T(1) + (ax * ax * v.x * v.x + ay * ay * v.y * v.y) / (v.z * v.z)), F(-0.5));", "return M::fma(F(1.0), M::sqrt(T(1) + (ax * ax * v.x * v.x + ay * ay * v.y * v.y) / (v.z * v.z)), F(-1.0));"),
}
only = sys.argv[1:]
work = tempfile.mkdtemp(prefix="oracle_mutants_")
missed = []
for name, (a, b) in MUT.items():
    if only and name not in only: continue
    assert src.count(a) >= 1, name
    tag = re.sub(r"\W+", "_", name)
    cpp = os.path.join(work, f"{tag}.cpp")
    open(cpp, "w").write(src.replace(a, b, 1))
    so = os.path.join(work, f"{tag}.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-march=x86-64-v3", "-fopenmp", "-w",
                    "-I", f"{ROOT}/oracle", "-shared", "-o", so, cpp], check=True)
    r = subprocess.run([sys.executable, "-m", "pytest", f"{ROOT}/tests/test_oracle_pins.py", f"{ROOT}/tests/test_oracle_closed_forms.py", "-q", "-p", "no:cacheprovider"],
                       env={**os.environ, "ORACLE_LIB": so}, capture_output=True, text=True, cwd=ROOT)
    failed = [l.split("::")[1].split(" ")[0] for l in r.stdout.splitlines() if l.startswith("FAILED")]
    print(f"{name}: {'caught' if failed else 'NOT CAUGHT'} by {sorted(set(f.split('[')[0] for f in failed))}", flush=True)
    if not failed:
        missed.append(name)
sys.exit(1 if missed else 0)
