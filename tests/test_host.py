"""Host side of the product (no GPU): the C ABI library loads and exports
every declared entry point; SC's input formats parse as the reference does."""
import json
import os
import re

import numpy as np
import pytest

import sc_amd
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sc_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = sc_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    # the Python binding covers the same set
    assert set(syms) <= set(sc_amd.SIGNATURES)


def test_binding_constants_match_the_header():
    txt = open(os.path.join(ROOT, "include", "sc_gpu.h")).read()
    m = re.search(r"#define SC_NOALLOC_MAX_PAIRS (\d+)", txt)
    assert m and int(m.group(1)) == sc_amd.NOALLOC_MAX_PAIRS


def test_topology_parser_matches_reference_params():
    g = json.load(open(os.path.join(GOLD, "topo_golden.json")))
    h = sc_amd.GpuSystem(-1)
    h.load_topology(g["top"])
    assert h.maxcut == pytest.approx(g["maxcut"], rel=1e-15)
    for key, row in g["rows"].items():
        i, j = (int(x) for x in key.split(","))
        np.testing.assert_allclose(h.pair_params(i, j), row, rtol=1e-14, atol=1e-15, err_msg=key)
    assert list(h.particle_types()) == list(range(1, 13))


def test_topology_errors_are_reported():
    h = sc_amd.GpuSystem(-1)
    with pytest.raises(sc_amd.SCError, match="wrong number of parameters"):
        h.load_topology("[Types]\nA 1 PSC 1.0 1.0 1.0\n[Molecules]\nA: {\nparticles: 1\n}\n[System]\nA 2\n")
    with pytest.raises(sc_amd.SCError, match="invalid keyword"):
        h.load_topology("[Typos]\nA 1 SPN 1.0 1.0\n")
    with pytest.raises(sc_amd.SCError, match="not defined"):
        h.load_topology("[Types]\nA 1 SPN 1.0 1.0\n[Molecules]\nA: {\nparticles: 1\n}\n[System]\nB 2\n")


def test_options_parser_postprocessing():
    from sc_amd import workloads as W
    h = sc_amd.GpuSystem(-1)
    h.load_options(W.options_text(transmx=0.15, rotmx=20.0))
    assert h.option("transmx") == pytest.approx(0.3)            # sim.h:365 doubles it
    assert h.option("rot_half") == pytest.approx(20.0 / 180 * np.pi / 2 * 0.5)   # sim.h:360
    h.load_options(W.options_text().replace("chainmmx = 0.0", "chainmmx = 0.4").replace("chainrmx = 0.0",
                                                                                          "chainrmx = 30.0"))
    assert h.option("chainmmx") == pytest.approx(0.8)           # sim.h:366
    assert h.option("chain_angle") == pytest.approx(30.0 / 180 * np.pi / 2)       # sim.h:362
    with pytest.raises(sc_amd.SCError, match="chains must be in range"):
        h.load_options(W.options_text().replace("chainrmx = 0.0", "chainrmx = 200.0"))
    bad = W.options_text().replace("seed = 12345\n", "")
    with pytest.raises(sc_amd.SCError, match="seed"):
        h.load_options(bad)


def test_rsa_configuration_has_no_overlap():
    from sc_amd import workloads as W
    w = W.workload("c2_cpsc_32768", n_override=300)
    h = sc_amd.GpuSystem(-1)
    h.load_topology(w["top"])
    pos, d, q = h.generate_rsa(w["types"], w["box"], seed=5)
    # closest segment distances with the oracle's restatement of minDistSegments
    n = len(pos)
    hl = 1.5
    mind = 1e9
    for i in range(n):
        for j in range(i + 1, n):
            dd = pos[i] - pos[j]
            r = w["box"] * (dd - np.rint(dd))
            if r @ r > (3 + 1.0) ** 2:
                continue
            v = po.min_dist_segments(d[i], d[j], hl, hl, r)
            mind = min(mind, np.linalg.norm(v))
    assert mind >= 1.0 - 1e-12
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0)
    assert np.allclose((d * q).sum(1), 0.0, atol=1e-12)


def test_cpp_dropin_adapter_compiles_and_links(tmp_path):
    """include/sc_gpu_totale.hpp (the TotalEnergyCalculator replacement)
    compiles as C++11 against the reference's member layout and links the
    library; without a GPU the constructor reports the missing device."""
    import subprocess
    exe = tmp_path / "totale_check"
    src = os.path.join(ROOT, "tests", "native", "totale_compile.cpp")
    r = subprocess.run(["g++", "-std=c++11", "-I" + os.path.join(ROOT, "include"), src,
                        "-L" + os.path.join(ROOT, "sc_amd", "lib"), "-lscamd",
                        "-Wl,-rpath," + os.path.join(ROOT, "sc_amd", "lib"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert "context" in r.stdout


def test_philox_known_answers(tmp_path):
    """The device generator (sc_rng.h, built for the host) reproduces the
    published Philox4x32-10 known-answer vectors of Random123 and its uniform
    conversions map the ends of their input ranges into [0, 1)."""
    import subprocess
    exe = tmp_path / "philox_kat"
    src = os.path.join(ROOT, "tests", "native", "philox_kat.cpp")
    r = subprocess.run(["g++", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("ok ") == 5


@pytest.mark.skipif(not os.path.exists("/root/reference/scOOP/mc/totalenergycalculator.h"),
                    reason="reference sources absent")
def test_reference_program_builds_with_the_gpu_backend_at_its_typedef():
    """The reference's whole program (scOOP 4.0, unmodified sources) compiles
    and links with GpuTotalE<Sim, Conf> as its TotalEnergyCalculator
    (mc/totalenergycalculator.h:1227), i.e. against the real Particle, Conf,
    Molecule, EMResize and the calls of Updater, MoveCreator and
    ClusterSampler (oracle/Makefile, target dropin)."""
    import subprocess
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8", "dropin"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    hdr = open(os.path.join(ROOT, "oracle", "_ref", "dropin", "src", "mc", "totalenergycalculator.h")).read()
    assert "typedef GpuTotalE<Sim, Conf> TotalEnergyCalculator;" in hdr
    assert "typedef TotalEMatrix<PairE> TotalEnergyCalculator;" not in hdr
    ref = open("/root/reference/scOOP/mc/totalenergycalculator.h").read()
    # everything else is the reference's text, line for line
    a = [ln for ln in ref.split("\n") if "typedef TotalEMatrix<PairE> TotalEnergyCalculator;" not in ln]
    b = [ln for ln in hdr.split("\n") if "GpuTotalE" not in ln and "sc_gpu_totale.hpp" not in ln]
    assert a == b
    assert os.access(os.path.join(ROOT, "oracle", "_ref", "scdropin"), os.X_OK)
