"""Write tests/golden/mixbox_latents.json: the 7-float Mixbox latents of the
palette colours the synthetic scenes use.

The latents come from the reference's own Mixbox (raytracing/src/mixbox/mixbox.cpp,
`mixbox_rgb_to_latent`, mixbox.h:74), compiled from its sources by
`make -C oracle ref` into oracle/_ref/libmixbox_ref.so.  Only the numbers are
committed; Mixbox's LUT is not copied.  Run in a container that has the
reference mounted:  python tests/golden/make_mixbox_latents.py
"""
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libmixbox_ref.so")

# Palette names/values as listed in raytracing/src/mixbox/mixbox.h:30-42.
PALETTE = {
    "cobalt_blue": (0, 33, 133),
    "hansa_yellow": (252, 211, 0),
    "cadmium_red": (255, 39, 2),
    "phthalo_green": (0, 60, 50),
    "ultramarine_blue": (25, 0, 89),
    "cadmium_orange": (255, 105, 0),
    "quinacridone_magenta": (128, 2, 46),
    "sap_green": (107, 148, 4),
    "burnt_sienna": (123, 72, 0),
    "cadmium_yellow": (254, 236, 0),
}


def main():
    if not os.path.exists(LIB):
        sys.exit("build the reference Mixbox first: make -C oracle ref")
    lib = ctypes.CDLL(LIB)
    lib.mixbox_rgb_to_latent.argtypes = [ctypes.c_ubyte] * 3 + [ctypes.POINTER(ctypes.c_float)]
    lib.mixbox_latent_to_rgb.argtypes = [ctypes.POINTER(ctypes.c_float)] + [ctypes.POINTER(ctypes.c_ubyte)] * 3
    out = {}
    for name, rgb in PALETTE.items():
        lat = (ctypes.c_float * 7)()
        lib.mixbox_rgb_to_latent(*rgb, lat)
        r, g, b = ctypes.c_ubyte(), ctypes.c_ubyte(), ctypes.c_ubyte()
        lib.mixbox_latent_to_rgb(lat, ctypes.byref(r), ctypes.byref(g), ctypes.byref(b))
        out[name] = {"srgb": list(rgb), "latent": [float(v) for v in lat],
                     "roundtrip_srgb": [r.value, g.value, b.value]}
    # a 50/50 mix of blue and yellow, decoded by the reference (a mixing fixture)
    zb = (ctypes.c_float * 7)(*out["cobalt_blue"]["latent"])
    zy = (ctypes.c_float * 7)(*out["hansa_yellow"]["latent"])
    zm = (ctypes.c_float * 7)(*[0.5 * a + 0.5 * b for a, b in zip(zb, zy)])
    r, g, b = ctypes.c_ubyte(), ctypes.c_ubyte(), ctypes.c_ubyte()
    lib.mixbox_latent_to_rgb(zm, ctypes.byref(r), ctypes.byref(g), ctypes.byref(b))
    doc = {"source": "reference Mixbox mixbox_rgb_to_latent (raytracing/src/mixbox/mixbox.cpp)",
           "colors": out,
           "mix_blue_yellow_50_50_srgb": [r.value, g.value, b.value]}
    with open(os.path.join(HERE, "mixbox_latents.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc["colors"]["cobalt_blue"]))


if __name__ == "__main__":
    main()
