"""CPU restatement (numpy) of the two passes after PLTPT in the reference's render graph:
AccumulatePass (Source/RenderPasses/AccumulatePass/AccumulatePass.cpp, Accumulate.cs.slang) and
ToneMapper (Source/RenderPasses/ToneMapper/ToneMapper.cpp, Luminance.ps.slang,
ToneMapping.ps.slang, Source/Falcor/Utils/Color/ColorUtils.h).

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels. Nothing in the library imports it.

Accumulation is restated operation by operation in IEEE float32 / float64 (numpy rounds every
operation and never fuses), so the device kernels must match it bit for bit. Tone mapping is
restated in float64 and compared within a tolerance: the reference's shader arithmetic is float
with hardware texture filtering, whose rounding neither side reproduces.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32

DOUBLE, SINGLE, SINGLE_COMPENSATED = 0, 1, 2
STOP, RESET, EMA = 0, 1, 2


class Accumulator:
    """AccumulatePass state machine (AccumulatePass.cpp:144-301) and its three shaders
    (Accumulate.cs.slang:74-198)."""

    def __init__(self, precision=SINGLE, max_frames=0, overflow=STOP, enabled=True, auto_reset=True):
        self.precision, self.max_frames, self.overflow = precision, max_frames, overflow
        self.enabled, self.auto_reset = enabled, auto_reset
        self.frame_count = 0
        self.dim = (0, 0)
        self.sum = self.corr = None

    def set_enabled(self, enabled):  # AccumulatePass.cpp:370-377
        if enabled != self.enabled:
            self.enabled = enabled
            self.frame_count = 0

    def execute(self, src, refresh=False):
        """src: [h, w, 4] float32 / uint32 / int32. Returns the output frame, or None when the pass
        stops at its frame limit (the output is then left as it was)."""
        if self.auto_reset and refresh:  # :146-174
            self.frame_count = 0
        if self.max_frames > 0 and self.frame_count == self.max_frames:  # :176-189
            if self.overflow == STOP:
                return None
            if self.overflow == RESET:
                self.frame_count = 0
        h, w = src.shape[:2]
        if (h, w) != self.dim:  # :200-205
            self.dim = (h, w)
            self.frame_count = 0
        cur = src.astype(F32)  # Texture2D<uint4/int4> read into float4
        integer = src.dtype != np.float32
        if not self.enabled and not integer:  # blit, :222-226
            return cur.copy()
        # accumulate(): clear when the count is 0 (:403-407), then the shader with gAccumCount.
        if self.frame_count == 0 or self.sum is None or self.sum.shape[:2] != (h, w):
            self.sum = np.zeros((h, w, 4), np.float64 if self.precision == DOUBLE else F32)
            self.corr = np.zeros((h, w, 4), F32)
        count = self.frame_count
        ema = self.max_frames > 0
        if self.max_frames == 0 or self.precision == SINGLE_COMPENSATED or self.frame_count < self.max_frames:  # :290-293
            self.frame_count += 1
        if not self.enabled:
            return cur.copy()
        wf = F32(1.0) / F32(count + 1)  # float literal / uint: a float division (:84, :157)
        if self.precision == SINGLE:  # :84-102
            if ema:
                out = self.sum + wf * (cur - self.sum)  # lerp(x, y, s) = x + s * (y - x)
                self.sum = out
            else:
                self.sum = self.sum + cur
                out = self.sum * wf
            return out
        if self.precision == SINGLE_COMPENSATED:  # :123-135
            y = cur - self.corr
            sn = self.sum + y
            out = sn / F32(count + 1)
            self.corr = (sn - self.sum) - y
            self.sum = sn
            return out
        wd = np.float64(wf)  # :157-188
        c64 = cur.astype(np.float64)
        if ema:
            self.sum = self.sum * (1.0 - wd) + c64 * wd
            return self.sum.astype(F32)
        self.sum = self.sum + c64
        return (self.sum * wd).astype(F32)


# ------------------------------------------------------------------------------ tone mapping

# Published colour-science matrices (ColorUtils.h:65-95): linear Rec.709 <-> CIE XYZ and CIECAM02's
# CAT02 chromatic adaptation with its inverse.
RGB_TO_XYZ = np.array([[0.4123907992659595, 0.3575843393838780, 0.1804807884018343],
                       [0.2126390058715104, 0.7151686787677559, 0.0721923153607337],
                       [0.0193308187155918, 0.1191947797946259, 0.9505321522496608]])
XYZ_TO_RGB = np.array([[3.2409699419045213, -1.5373831775700935, -0.4986107602930033],
                       [-0.9692436362808798, 1.8759675015077206, 0.0415550574071756],
                       [0.0556300796969936, -0.2039769588889765, 1.0569715142428784]])
XYZ_TO_LMS_CAT02 = np.array([[0.7328, 0.4296, -0.1624], [-0.7036, 1.6975, 0.0061], [0.0030, 0.0136, 0.9834]])
LMS_TO_XYZ_CAT02 = np.array([[1.096123820835514, -0.278869000218287, 0.182745179382773],
                             [0.454369041975359, 0.473533154307412, 0.072097803717229],
                             [-0.009627608738429, -0.005698031216113, 1.015325639954543]])
LUMA = np.array([0.299, 0.587, 0.114])  # Luminance.ps.slang:34-37, ToneMapping.ps.slang:46-49


def color_temperature_xyz(T):
    """XYZ (Y = 1) of a blackbody at T K: Kang et al. 2002 Planckian-locus fit (ColorUtils.h:144-187)."""
    t = float(T)
    if T < 4000:
        xc = -0.2661239e9 / t ** 3 - 0.2343580e6 / t ** 2 + 0.8776956e3 / t + 0.179910
    else:
        xc = -3.0258469e9 / t ** 3 + 2.1070379e6 / t ** 2 + 0.2226347e3 / t + 0.240390
    if T < 2222:
        yc = -1.1063814 * xc ** 3 - 1.34811020 * xc ** 2 + 2.18555832 * xc - 0.20219683
    elif T < 4000:
        yc = -0.9549476 * xc ** 3 - 1.37418593 * xc ** 2 + 2.09137015 * xc - 0.16748867
    else:
        yc = 3.0817580 * xc ** 3 - 5.87338670 * xc ** 2 + 3.75112997 * xc - 0.37001483
    return np.array([xc / yc, 1.0, (1.0 - xc - yc) / yc])


def white_balance(T):
    """Von Kries adaptation from temperature T to 6500 K in CAT02 LMS, in Rec.709 RGB (ColorUtils.h:202-218)."""
    ma = XYZ_TO_LMS_CAT02 @ RGB_TO_XYZ
    inv_ma = XYZ_TO_RGB @ LMS_TO_XYZ_CAT02
    wd = XYZ_TO_LMS_CAT02 @ color_temperature_xyz(6500.0)
    ws = XYZ_TO_LMS_CAT02 @ color_temperature_xyz(T)
    return inv_ma @ np.diag(wd / ws) @ ma


def color_transform(o):
    """ToneMapper::updateColorTransform (ToneMapper.cpp:535-547); `o` holds the pass options."""
    wb = white_balance(o["whitePoint"]) if o["whiteBalance"] else np.eye(3)
    manual = 1.0 if o["autoExposure"] else (0.01 * o["filmSpeed"]) / (o["shutter"] * o["fNumber"] ** 2)
    return wb * 2.0 ** o["exposureCompensation"] * manual


def bit_floor(v):
    b = 1
    while b * 2 <= v:
        b *= 2
    return b


def luminance_image(img):
    """The luminance pass (ToneMapper.cpp:230-241, Luminance.ps.slang:39-42): bit_floor(W) x
    bit_floor(H) texels, each log2(max(1e-4, luma)) of a bilinear sample of the source at the
    texel centre."""
    img = np.asarray(img, np.float64)
    H, W = img.shape[:2]
    h, w = bit_floor(H), bit_floor(W)
    u = (np.arange(w) + 0.5) * (W / w) - 0.5
    v = (np.arange(h) + 0.5) * (H / h) - 0.5
    x0 = np.floor(u).astype(int)
    y0 = np.floor(v).astype(int)
    fu = (u - x0)[None, :, None]
    fv = (v - y0)[:, None, None]
    x1 = np.minimum(x0 + 1, W - 1)
    y1 = np.minimum(y0 + 1, H - 1)
    rgb = img[..., :3]
    top = rgb[y0][:, x0] * (1 - fu) + rgb[y0][:, x1] * fu
    bot = rgb[y1][:, x0] * (1 - fu) + rgb[y1][:, x1] * fu
    s = top * (1 - fv) + bot * fv
    return np.log2(np.maximum(1e-4, s @ LUMA))


def mip_top(lum):
    """The last level of the 2x2-box mip chain (Texture::generateMips, Core/API/Texture.cpp:510-540)."""
    a = np.asarray(lum, np.float64)
    while a.shape != (1, 1):
        if a.shape[0] > 1:
            a = 0.5 * (a[0::2] + a[1::2])
        if a.shape[1] > 1:
            a = 0.5 * (a[:, 0::2] + a[:, 1::2])
    return float(a[0, 0])


def uc2(x):
    """Uncharted 2 filmic curve (ToneMapping.ps.slang:85-96)."""
    A, B, C, D, E, F = 0.22, 0.3, 0.1, 0.2, 0.01, 0.3
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F


def tone_map(img, o, mean_log_luminance=None, transform=None):
    """The tone-mapping pass (ToneMapping.ps.slang:145-168) in float64; returns [h, w, 4].
    `transform` replaces color_transform(o) (e.g. by the float matrix a device run used)."""
    img = np.asarray(img, np.float64)
    c = img[..., :3].copy()
    if o["autoExposure"]:
        if mean_log_luminance is None:
            mean_log_luminance = mip_top(luminance_image(img))
        c *= 0.042 / 2.0 ** mean_log_luminance
    c = c @ (color_transform(o) if transform is None else np.asarray(transform, np.float64)).T
    op = int(o["operator"])
    with np.errstate(divide="ignore", invalid="ignore"):
        if op == 1:  # Reinhard (:56-62)
            L = (c @ LUMA)[..., None]
            c = c * ((L / (L + 1)) / L)
        elif op == 2:  # ReinhardModified as the reference writes it (:65-70)
            L = (c @ LUMA)[..., None]
            wl = o["whiteMaxLuminance"]
            c = c * ((L * (1 + L / (wl * wl)) * (1 + L)) / L)
        elif op == 3:  # HejiHableAlu (:74-81)
            c = np.maximum(0.0, c - 0.004)
            c = (c * (6.2 * c + 0.5)) / (c * (6.2 * c + 1.7) + 0.06)
            c = c ** 2.2
        elif op == 4:  # HableUc2 (:98-105)
            c = uc2(2.0 * c) * (1.0 / uc2(o["whiteScale"]))
        elif op == 5:  # Aces (:108-121)
            c = c * 0.6
            c = np.clip((c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14), 0.0, 1.0)
    if o["clamp"]:
        c = np.where(np.isnan(c), 0.0, np.clip(c, 0.0, 1.0))  # saturate(NaN) = 0
    out = np.empty(img.shape, np.float64)
    out[..., :3] = c
    out[..., 3] = img[..., 3]
    return out


TONEMAP_DEFAULTS = dict(exposureCompensation=0.0, autoExposure=False, filmSpeed=100.0, whiteBalance=False,
                        whitePoint=6500.0, operator=4, clamp=True, whiteMaxLuminance=1.0, whiteScale=11.2,
                        fNumber=1.0, shutter=1.0)
