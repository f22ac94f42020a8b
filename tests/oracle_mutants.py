"""Mutation check of the oracle's pins (CPU only, not collected by pytest).

Plants one plausible transcription mistake at a time in a scratch copy of the
oracle sources, builds that copy with g++ into a temporary directory, and runs
the oracle's CPU tests against it (through WR_ORACLE_LIB). Every mutant must
make at least one test fail; the script exits non-zero if one survives.

    python tests/oracle_mutants.py          (about 5 minutes on 8 cores)
"""
import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

R = str(Path(__file__).resolve().parent.parent)
W = tempfile.mkdtemp(prefix="oracle_mutants_")
M = [
 ("rr no divide", "newThr = mk<T>(thr.x / q, thr.y / q, thr.z / q)", "newThr = mk<T>(thr.x, thr.y, thr.z)"),
 ("rr lum permuted", "T lum = T(0.2126f) * thr.x + T(0.7152f) * thr.y + T(0.0722f) * thr.z;", "T lum = T(0.0722f) * thr.x + T(0.7152f) * thr.y + T(0.2126f) * thr.z;"),
 ("rr depth factor", "(T(depth) - rrStart)", "T(depth)"),
 ("rr reversed", "if (xi < q) { newThr", "if (xi > q) { newThr"),
 ("rr minprob", "T minProb = isDelta ? T(0.5f) : T(0.05f);", "T minProb = isDelta ? T(0.5f) : T(0.1f);"),
 ("asym inverted", "return cg / cs;", "return cs / cg;"),
 ("refract sign", "T k = eta * cosI - cosT;", "T k = eta * cosI + cosT;"),
 ("glass corr on gn", "T corr = asymCorrection(rf, hit.normal, hit.geometricNormal);", "T corr = asymCorrection(rf, hit.geometricNormal, hit.normal);"),
 ("nmap t/b swap", "V3<T> w = mk<T>(tg.x * n.x + bt.x * n.y + hit.normal.x * n.z,", "V3<T> w = mk<T>(bt.x * n.x + tg.x * n.y + hit.normal.x * n.z,"),
 ("uv swap", "hit.u = ub * a0 + vb * a1 + wb * a2;", "hit.u = ub * a0 + wb * a1 + vb * a2;"),
 ("normal interp swap", "sN = normalize(ub * n0 + vb * n1 + wb * n2);", "sN = normalize(vb * n0 + ub * n1 + wb * n2);"),
 ("bilinear fx", "T top = tx(i0, j0, c) * (T(1) - fx) + tx(i1, j0, c) * fx;", "T top = tx(i0, j0, c) * fx + tx(i1, j0, c) * (T(1) - fx);"),
 ("bilinear no half", "T x = uf * T(tex.w) - T(0.5f);", "T x = uf * T(tex.w);"),
 ("schlick mix", "V3<T> F0 = mk<T>(f0 * (T(1) - metal) + albedo.x * metal,", "V3<T> F0 = mk<T>(f0 * metal + albedo.x * (T(1) - metal),"),
 ("kD no metal", "V3<T> kD = mk<T>((T(1) - F.x) * (T(1) - metal),", "V3<T> kD = mk<T>((T(1) - F.x),"),
 ("ggx weight no cos", "ray.throughput = mk<T>(ray.throughput.x * brdf.x * cosTheta / pdf,", "ray.throughput = mk<T>(ray.throughput.x * brdf.x / pdf,"),
 ("camera y flip", "T ndcY = T(1) - T(2) * (T(float(py)) + r2) / T(float(p.height));", "T ndcY = T(2) * (T(float(py)) + r2) / T(float(p.height)) - T(1);"),
 ("camera no aspect", "T ndcX = (T(2) * (T(float(px)) + r1) / T(float(p.width)) - T(1)) * T(cam.aspect);", "T ndcX = (T(2) * (T(float(px)) + r1) / T(float(p.width)) - T(1));"),
 ("lens focal", "V3<T> fp = mk<T>(P.x + d.x * fd, P.y + d.y * fd, P.z + d.z * fd);", "V3<T> fp = mk<T>(P.x + d.x * fd, P.y + d.y * fd, P.z + d.z * fd * T(0.5f));"),
 ("lens radius", "T radius = sqrtT(r4) * T(cam.lensRadius);", "T radius = r4 * T(cam.lensRadius);"),
 ("clamp per channel", "return mk<T>(L.x * scale, L.y * scale, L.z * scale);", "return mk<T>(fminT(L.x, maxL), fminT(L.y, maxL), fminT(L.z, maxL));"),
 ("offset side", "V3<T> od = dot(dir, n) > T(0) ? n : -n;", "V3<T> od = dot(dir, n) >= T(0) ? n : -n;"),
 ("tri sample", "T b1 = su * (T(1) - u2);", "T b1 = (T(1) - u2);"),
 ("max bounces", "if (ray.depth >= p.maxBounces) terminateRay(ray);", "if (ray.depth > p.maxBounces) terminateRay(ray);"),
 ("mis weight", "w = powerHeuristic(ray.prevPdf, lightPdf);", "w = powerHeuristic(lightPdf, ray.prevPdf);"),
 ("fresnel", "return (rs * rs + rp * rp) * T(0.5f);", "return (rs * rs + rp * rp) * T(0.45f);"),
]
src = open(R + "/oracle/wr_oracle.cpp").read()
hdr = R + "/oracle/wr_oracle.hpp"
alive = []
for name, a, b in M:
    if src.count(a) != 1:
        print("PATTERN?", name, src.count(a)); alive.append(name); continue
    d = W + "/src"
    os.makedirs(d, exist_ok=True)
    open(d + "/wr_oracle.cpp", "w").write(src.replace(a, b))
    shutil.copy(hdr, d + "/wr_oracle.hpp")
    shutil.copy(R + "/oracle/oracle_capi.cpp", d + "/oracle_capi.cpp")
    so = W + "/liboracle_mut.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-shared", "-I", R + "/oracle",
                    "-o", so, d + "/wr_oracle.cpp", d + "/oracle_capi.cpp"], check=True,
                   cwd=d)
    env = dict(os.environ, WR_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "tests/test_oracle_pins.py", "tests/test_oracle_math.py", "tests/test_oracle_paths.py"],
                       cwd=R, env=env, capture_output=True, text=True)
    killed = r.returncode != 0
    first = [l for l in r.stdout.splitlines() if l.startswith("FAILED")][:1]
    print(("killed " if killed else "ALIVE  ") + name, first)
    if not killed: alive.append(name)
print("alive:", alive)
shutil.rmtree(W, ignore_errors=True)
sys.exit(1 if alive else 0)
