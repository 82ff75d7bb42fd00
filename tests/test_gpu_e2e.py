"""End-to-end optimisation at reduced size (tests/e2e_train.py at 64^3, 8 training views and
2 held-out views of 64x64, 4096 rays x 128 samples per iteration, 150 iterations, no
progressive growing): the fused forward of the ground-truth grid matches the FP64 oracle's
images, training raises the held-out PSNR, the trained grid's loss matches an FP64 evaluation,
and BF16 storage trains to within 0.1 dB of the FP32 result's held-out PSNR."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import e2e_train  # noqa: E402
from util import say  # noqa: E402

pytestmark = pytest.mark.gpu


def test_e2e_small_bf16_and_fp32(tmp_path):
    say("e2e 64^3, 150 iterations, bf16 + fp32")
    r = e2e_train.main(["--grid", "64", "--views", "8", "--heldout", "2", "--res", "64", "--iters", "150",
                        "--rays", "4096", "--per-camera", "1", "--samples", "128", "--progressive", "0", "--plain-steps", "0",
                        "--log-every", "50", "--threads", "4", "--out", str(tmp_path / "e2e.json")])
    par = r["gt_forward_parity"]
    # FP32 kernel vs FP64 oracle on the same grid and rays; the t_stop cut can differ only where
    # the transmittance is within rounding of 1e-4, which moves a pixel by less than 1e-4
    assert par["max_abs_err"] <= 1e-4, par
    for name in ("bfloat16", "float32"):
        f = r["fused"][name]
        say(f"  {name}: held-out PSNR {f['heldout_psnr_rgb']} dB (init {f['heldout_psnr_rgb_init']}), "
            f"final loss {f['final_loss_eval']}")
        assert f["heldout_psnr_rgb"] >= f["heldout_psnr_rgb_init"] + 3.0, f
        assert f["loss_history"][-1][1] < f["loss_history"][0][1], f
    # the trained grid's loss by the fused kernel vs the FP64 oracle on the FP32 master: FP32
    # storage to the forward bar; BF16 storage carries the rounding of the render copy
    assert r["fused"]["float32"]["final_loss_eval"]["rel_err"] <= 1e-4
    assert r["fused"]["bfloat16"]["final_loss_eval"]["rel_err"] <= 1e-2
    say("  held-out PSNR delta bf16 - fp32:", r["heldout_psnr_delta_bf16_vs_fp32_db"], "dB")
    assert abs(r["heldout_psnr_delta_bf16_vs_fp32_db"]) <= 0.1, r["heldout_psnr_delta_bf16_vs_fp32_db"]
