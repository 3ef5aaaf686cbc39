"""The small scenes of the image-parity tests, regenerated from fixed seeds.

spheres          64 UV spheres at 64x64, constant environment 0.5
spheres_sky_tex  24 spheres at 64x48, textured ground, HDR sky with a sun lobe
terrain_bokeh    64x32-cell heightfield at 96x54, 4 textured materials,
                 4 point lights, HDR sky, defocus
"""
import numpy as np

import eleven
from eleven import scenegen

SCENES = ["spheres", "spheres_sky_tex", "terrain_bokeh"]
DEPTH = 5


def make_scene(kind, root):
    if kind == "spheres":
        d = scenegen.sphere_field(root / kind, res=(64, 64), seed=0)
    elif kind == "spheres_sky_tex":
        rng = np.random.default_rng(4)
        tex = (rng.uniform(0, 1, size=(32, 32, 3)) * 255).astype(np.uint8)
        d = scenegen.sphere_field(root / kind, n_spheres=24, res=(64, 48), seed=1, ground_texture=tex,
                                  hdri_sky=scenegen.sky_hdr(128, 64, peak=50.0, sigma_deg=6.0))
    elif kind == "terrain_bokeh":
        d = scenegen.heightfield(root / kind, cells=(64, 32), extent=(12.8, 6.4), res=(96, 54), seed=2,
                                 n_mats=4, tex_size=64, lights=4, sky=(128, 64), aperture=0.05, bokeh=True)
    else:
        raise ValueError(kind)
    return eleven.SceneData.load(d)
