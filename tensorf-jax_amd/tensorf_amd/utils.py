"""utils.py of the reference (utils.py:4-9)."""
import math


def psnr_from_mse(mse) -> float:
    """-10 log10(max(mse, 1e-10)) (utils.py:4-9)."""
    return -10.0 * math.log10(max(float(mse), 1e-10))
