"""The C++ drop-in classes (plbpd_amd/host) through tests/cpp/test_dropin and
the plbpd_run runner."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_dropin")
BIN64 = os.path.join(ROOT, "tests", "cpp", "_build", "test_dropin_f64")
RUN = os.path.join(ROOT, "plbpd_amd", "lib", "plbpd_run")

needs_bin = pytest.mark.skipif(not os.path.exists(BIN), reason="tests/cpp not built (run __graft_entry__.build())")


@needs_bin
@pytest.mark.parametrize("binary", [BIN, BIN64], ids=["float_fields", "double_fields"])
def test_dropin_reference_known_answers(binary):
    r = subprocess.run([binary, "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@needs_bin
@pytest.mark.gpu
@pytest.mark.parametrize("binary", [BIN, BIN64], ids=["float_fields", "double_fields"])
def test_dropin_march_in_time_matches_oracle(binary):
    """The drop-in classes with the reference's float host fields (the
    default build) and with double fields (-DPLBPD_REAL=double)."""
    r = subprocess.run([binary, "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr


MAIN14 = os.path.join(ROOT, "tests", "cpp", "_build", "main_cxx14")


def test_reference_main_compiles_as_cxx14():
    """A caller of the reference's public interface (the call sequence of
    main.cpp:13-31) compiles as C++14 (the reference's standard,
    src/CMakeLists.txt:6) against the drop-in headers."""
    src = os.path.join(ROOT, "tests", "cpp", "main_cxx14.cpp")
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", src,
                        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "plbpd_amd", "host")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("real", ["double", "float"])
def test_reference_interface_signatures(real):
    """Every public member of the reference's headers exists in the drop-in
    with the same type (tests/cpp/api_cxx14.cpp, static assertions), under
    C++14, for double fields and for the reference's float fields."""
    src = os.path.join(ROOT, "tests", "cpp", "api_cxx14.cpp")
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", f"-DPLBPD_REAL={real}",
                        src, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "plbpd_amd", "host")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_reference_main_runs(tmp_path):
    """The same program on the GPU: Couette3D (200x50x50 D3Q19, 200 steps,
    literal, the default build's float fields and FP32 device state), writing
    the initial and the final state and printing the class timers, whose
    collide line reports GPU time."""
    r = subprocess.run([MAIN14], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Time taken by timeloop" in r.stdout
    init, final = (tmp_path / "Couette3D_begin.h5").read_bytes(), (tmp_path / "Couette3D_end.h5").read_bytes()
    assert len(init) == len(final) > 0 and init != final
    collide = [l for l in r.stdout.splitlines() if l.startswith("-collide:")]
    assert collide and float(collide[0].split(":")[1].strip().rstrip("s")) > 0, r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["Couette3D", "--steps", "20"],
                                  ["Poiseuille2D", "--steps", "200", "--fused", "--variant", "f32s"]])
def test_runner(tmp_path, args):
    out = tmp_path / "state.h5"
    r = subprocess.run([RUN, *args, "--out", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Time taken by timeloop" in r.stdout
    assert out.exists() and out.stat().st_size > 0


@pytest.mark.gpu
def test_runner_scenario_file(tmp_path):
    """plbpd_run with a scenario file instead of a preset name: lattice,
    boundary, storage, step semantics, initial state and solids from text."""
    cfg = tmp_path / "porous.cfg"
    cfg.write_text("lattice = D3Q19\ndimensions = 32 16 12\nviscosity = 0.05\nforce = 1e-5 0 0\n"
                   "boundary.east = periodic\nboundary.west = periodic\n"
                   "boundary.north = periodic\nboundary.south = periodic\n"
                   "boundary.up = noslip\nboundary.down = noslip\nvelocity.up = 0.01 0 0\n"
                   "timesteps = 40\nvariant = f16s\nmode = fused\ninit = perturbed 5 1e-3\n"
                   "solids.fraction = 0.05\nsolids.seed = 3\n")
    out = tmp_path / "state.h5"
    r = subprocess.run([RUN, str(cfg), "--out", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Time taken by timeloop" in r.stdout
    assert out.exists() and out.stat().st_size > 0


@pytest.mark.gpu
@pytest.mark.parametrize("args,stage", [(["Couette3D", "--steps", "30"], "device collide:"),
                                        (["Couette3D", "--steps", "30", "--fused"], "device fused:")])
def test_runner_device_times(tmp_path, args, stage):
    """With --device-times the runner switches the state's HIP-event timers on
    (DeviceConfig::device_timers) and prints the GPU time of each stage."""
    out = tmp_path / "state.h5"
    r = subprocess.run([RUN, *args, "--device-times", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith(stage)]
    assert line and float(line[0].split(":")[1].strip().rstrip("s")) > 0, r.stdout
