"""CPU restatement of the reference's frame post-processing (test
infrastructure only; imported by tests as the checker).

- AccumulationBuffer::Apply (Core/Postproc/AccumulationBuffer.h:59-88): mean of
  accumulated frames.
- ImageDevice::WriteRadianceBufferToDevice (Core/Device/ImageDevice.cpp:137-148):
  frame pixel (x, y) lands at (W-1-x, H-1-y).
- GlobalTone::Apply (Core/Postproc/GlobalTone.h:37-53): c / (c + 1).
- DragoTone::Apply (Core/Postproc/DragoTone.h:37-80): bias 0.8, scale 0.95,
  per-channel maximum Lmax over the region starting from 0, then
  0.95 / log10(Lmax + 1) * ln(L + 1) / ln(2 + 8 (L / Lmax)^(ln 0.8 / ln 0.5)),
  clamped to [0, 1].
- ImagePPM::Save (Core/Image/ImagePPM.h:85-110): bytes min(255, int(256 v)).
"""
import numpy as np


def resolve(accum, samples, flip=False):
    img = np.asarray(accum, np.float64) / samples
    return img[::-1, ::-1] if flip else img


def global_tone(img):
    img = np.asarray(img, np.float64)
    return img / (img + 1.0)


def drago_tone(img):
    img = np.asarray(img, np.float64)
    lmax = np.maximum(img.reshape(-1, 3).max(axis=0), 0.0)
    bias = np.log(0.8) / np.log(0.5)
    v = 0.95 / (np.log(lmax + 1.0) / np.log(10.0)) * np.log(img + 1.0) / np.log(2.0 + (img / lmax) ** bias * 8.0)
    return np.clip(v, 0.0, 1.0)


def ppm_bytes(img):
    v = (256.0 * np.asarray(img, np.float32)).astype(np.int64)   # C float-to-int conversion truncates
    return np.minimum(v, 255).astype(np.uint8)
