"""The per-pixel colour bar shared by the path-tracer parity tests (GPU pass against the oracle on
the same scene, frame and pixels): |GPU - oracle| <= 1e-3 |oracle| + a floor of 1e-4 x the image's
mean pixel value, per channel. Channels whose value is the rounding residue of a cancelled sum have
no relative accuracy; the floor is the image's own scale, not an absolute constant."""
import numpy as np

COLOR_REL, COLOR_FLOOR = 1e-3, 1e-4


def color_floor(o_col) -> float:
    return COLOR_FLOOR * max(float(np.abs(o_col[:, :3]).mean()), 1e-30)


def color_close(g_col, o_col):
    """Per pixel: every RGB channel within the bar."""
    floor = color_floor(o_col)
    return np.all(np.abs(g_col[:, :3] - o_col[:, :3]) <= COLOR_REL * np.abs(o_col[:, :3]) + floor, axis=1)
