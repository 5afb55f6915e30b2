"""Shared by the asset tests: the committed smoke subset and cameras framed on its bounds."""
import math
import os

from hdsr import assets, camera

SMOKE = os.path.join(os.path.dirname(__file__), "golden", "smoke_subset", "scaniverse-first-smoke.json")


def smoke_cameras(attrs, width=800, height=600):
    """Three orbit views around the declared centre at 0.45 of framingFromBounds' distance
    (fovY 60 degrees), where the capture fills most of the frame."""
    fr = assets.framingFromBounds(attrs.bounds)
    f = height / (2 * math.tan(math.pi / 6))
    out = []
    for az, el in [(0.0, 0.35), (2.1, 0.2), (4.2, 0.6)]:
        eye = camera.orbit_eye(0.45 * fr["distance"], az, el, target=fr["target"])
        out.append(camera.Camera(eye, fr["target"], f, f, width, height))
    return out
