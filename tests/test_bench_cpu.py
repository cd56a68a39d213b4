"""CPU checks of bench.py's contract pieces that need no GPU: the single-GPU
refusal for --gpus N > 1, and the roofline traffic read from the committed
counter summary."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_multi_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert p.returncode == 2
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_traffic_summary_parses_and_exceeds_algorithmic_bytes():
    b = _bench()
    for (config, kind), name in b.TRAFFIC_PROFILES.items():
        assert os.path.exists(os.path.join(ROOT, "profiles", name)), name
        alone = b.profile_value(config, kind, "alone_tflops")
        peak = 157.3 if kind.startswith("fp32") else 78.6
        assert alone is not None and 0 < alone < peak, (kind, alone)
        t = b.traffic_from_profiles(config, kind)
        if t is None:  # a stand-alone timing without counter runs
            continue
        cfg = b.CONFIGS[config]
        n = cfg["npy"] * cfg["npz"] * 4
        batch = cfg["per_step"] // 4  # energies per lane

        def update_bytes(cols, k):  # the m x k and k x cols operands read once, C read and written
            return batch * 8 * (n * k + k * cols + 2 * n * cols)
        if kind == "fp32x3_gj_update":
            # groups of four 128-column blocks (18 blocks: four groups of 512, one of 256):
            # per inverse 16 updates (512 x 128) and 2 (256 x 128) of a group's own columns,
            # 4 (1792 x 512) and 1 (2048 x 256) of the rest; the average launch
            shapes = [(16, 512, 128), (2, 256, 128), (4, n - 512, 512), (1, n - 256, 256)]
            alg = sum(c * update_bytes(cols, k) for c, cols, k in shapes) / sum(c for c, _, _ in shapes)
        elif "gj_update" in kind:
            alg = update_bytes(n, 128)  # rank-128 update over every column
        else:  # A and B read once, C written once
            alg = 3 * batch * n * n * (8 if kind.startswith("fp32") else 16)
        assert t >= alg, (t, alg)
    assert b.traffic_from_profiles("c4", "fp64_gemm_3m") is not None
    assert b.traffic_from_profiles("c4", "fp32_gj_update_3m") is not None
    assert b.traffic_from_profiles("c4", "fp32x3_gj_update") is not None
    assert b.traffic_from_profiles("c2", "fp64_gemm_3m") is None


def test_dominant_kind_is_largest_summed_duration():
    b = _bench()
    prof = {k: (0.0, 0.0, 0) for k in b.GEMM_KINDS}
    prof.update(fp64_gj_panel=(1e9, 500.0, 100), fp32_gj_panel=(0.0, 0.0, 0))
    prof["fp32_gemm_3m"] = (1e15, 900.0, 300)
    prof["fp64_gemm_3m"] = (1e15, 1800.0, 200)
    assert b.dominant_kind(prof) == "fp64_gemm_3m"
    t = b.kernel_table(prof, "c4")
    assert "gbytes_per_s" in t["fp64_gj_panel"] and "tflops" in t["fp64_gemm_3m"]
    assert "fp32_gj_panel" not in t


def test_profile_kinds_match_the_c_abi():
    # negf_profile_end fills one entry per kind; the Python buffers must be exactly
    # as long as the arrays the header documents (a shorter buffer is overwritten)
    import re
    from nonequilibriumgreensfunction_amd import capi
    with open(os.path.join(ROOT, "include", "negf_capi.h")) as f:
        text = f.read()
    m = re.search(r"each argument is an array of (\d+)", text)
    assert m and int(m.group(1)) == len(capi.PROFILE_KINDS)
    src = open(os.path.join(ROOT, "nonequilibriumgreensfunction_amd", "csrc", "negf_device.h")).read()
    k = re.search(r"constexpr int kTimerKinds = (\d+);", src)
    assert k and int(k.group(1)) == len(capi.PROFILE_KINDS)
    b = _bench()
    assert set(b.GEMM_KINDS) <= set(capi.PROFILE_KINDS)
