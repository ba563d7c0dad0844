"""Writes the .scene loader fixtures under tests/golden/scenes/ and their
expected outputs.

Inputs (written by this script): feature.scene and camera_matrix.scene in the
reference's .scene format, small OBJ meshes, PNG textures of every colour
type (RGBA, RGB, grey, palette with tRNS, 16-bit grey) and a Radiance HDR
environment map (run-length encoded scanlines).

Expected outputs: the reference's own loader (src/loaders/Loader.cpp,
compiled into oracle/_ref/libref_scene.so) + ProcessScene on those inputs,
stored as expected_<name>.npz. Run from the repository root:

    python tests/golden/make_scene_fixtures.py
"""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "scenes")
sys.path.insert(0, os.path.dirname(HERE))

BOX_OBJ = """# unit box with uvs and normals, quads (fan-triangulated)
v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
v -1 -1 1
v 1 -1 1
v 1 1 1
v -1 1 1
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 -1
vn 0 0 1
vn 0 -1 0
vn 0 1 0
vn -1 0 0
vn 1 0 0
f 1/1/1 4/4/1 3/3/1 2/2/1
f 5/1/2 6/2/2 7/3/2 8/4/2
f 1/1/3 2/2/3 6/3/3 5/4/3
f 4/1/4 8/2/4 7/3/4 3/4/4
f 1/1/5 5/2/5 8/3/5 4/4/5
f 2/1/6 3/2/6 7/3/6 6/4/6
"""

TRI_OBJ = """# a plane of two triangles, positions and one normal (no uvs)
v -4 0 -4
v 4 0 -4
v 4 0 4
v -4 0 4
vn 0 1 0
f 1//1 3//1 2//1
f 1//1 4//1 3//1
"""

FEATURE_SCENE = """# every block and key of the format
renderer
{
	resolution 320 240
	maxdepth 6
	maxspp 64
	tilewidth 80
	tileheight 60
	envmapfile sky.hdr
	envmapintensity 1.75
	envmaprotation 45
	enablerr false
	rrdepth 3
	enabletonemap true
	enableaces true
	texarraywidth 16
	texarrayheight 16
	openglnormalmap false
	hideemitters true
	enablebackground true
	transparentbackground false
	backgroundcolor 0.1 0.2 0.3
	independentrendersize true
	windowresolution 640 480
	enableroughnessmollification true
	roughnessmollificationamt 0.25
	enablevolumemis true
	enableuniformlight true
	uniformlightcolor 0.5 0.4 0.3
}

camera
{
	position 0 3 9
	lookat 0 1 0
	fov 50
	aperture 0.05
	focaldist 8.5
}

material red
{
	color 0.8 0.1 0.1
	roughness 0.3
	metallic 0.2
	subsurface 0.1
	speculartint 0.4
	anisotropic 0.3
	sheen 0.5
	sheentint 0.6
	clearcoat 0.7
	clearcoatgloss 0.8
	ior 1.33
	albedotexture rgba8.png
	metallicroughnesstexture rgb8.png
	normaltexture grey8.png
	emissiontexture palette.png
}

material glass
{
	color 1 1 1
	spectrans 1.0
	roughness 0.0
	mediumtype scatter
	mediumdensity 2.5
	mediumcolor 0.9 0.4 0.1
	mediumanisotropy 0.3
	alphamode blend
	opacity 0.0
}

material smoke
{
	mediumtype absorb
	mediumdensity 0.5
	alphamode mask
	alphacutoff 0.4
	albedotexture grey16.png
}

material glow
{
	emission 4 3 2
	mediumtype emissive
	alphamode opaque
	albedotexture rgba8.png
}

material red
{
	color 0 1 0
}

light
{
	type quad
	position -1 5 -1
	v1 1 5 -1
	v2 -1 5 1
	emission 10 9 8
}

light
{
	type sphere
	position 2 4 0
	radius 0.5
	emission 20 20 20
}

light
{
	type distant
	position 0.3 1 0.2
	emission 2 2 2
}

mesh
{
	name floor plane
	file tri.obj
	material red
}

mesh
{
	file box.obj
	material glass
	position 1 1 0
	scale 0.5 0.75 0.5
	rotation 0 0.3826834 0 0.9238795
}

mesh
{
	file box.obj
	material nosuchmaterial
	position -2 1 0
}

mesh
{
	file box.obj
	material smoke
	matrix 1 0 0 0 0 1 0 0 0 0 1 0 0 2 -3 1
}

mesh
{
	file box.obj

	material glow
	position 3 3 3
}

mesh
{
	file missing.obj
	material red
}
"""

CAMERA_MATRIX_SCENE = """renderer
{
	resolution 200 100
	enableuniformlight true
	maxdepth 3
}

camera
{
	fov 35
	matrix 1 0 -1 5 0 1 -0.2 1 0 0 -0.4 2 0 0 0 1
}

material grey
{
	color 0.5 0.5 0.5
}

mesh
{
	file box.obj
	material grey
}
"""


def write_png_files():
    from PIL import Image
    rng = np.random.default_rng(0x5CE4E)
    rgba = rng.integers(0, 256, (16, 16, 4), dtype=np.uint8)
    Image.fromarray(rgba, "RGBA").save(os.path.join(OUT, "rgba8.png"))
    Image.fromarray(rgba[..., :3].copy(), "RGB").save(os.path.join(OUT, "rgb8.png"))
    Image.fromarray(rgba[..., 0].copy(), "L").save(os.path.join(OUT, "grey8.png"))
    pal = Image.fromarray((np.arange(256).reshape(16, 16) % 7).astype(np.uint8), "P")
    pal.putpalette([v for i in range(7) for v in (30 * i, 255 - 30 * i, 17 * i)])
    pal.save(os.path.join(OUT, "palette.png"), transparency=bytes([255, 0, 128, 255, 64, 255, 255]))
    g16 = (np.arange(256).reshape(16, 16).astype(np.uint16) * 251)
    Image.fromarray(g16.astype(np.int32), "I").convert("I;16").save(os.path.join(OUT, "grey16.png"))


def write_hdr(path, img):
    """Radiance RGBE with new-style run-length encoded scanlines."""
    h, w, _ = img.shape
    m = img.max(axis=2)
    e = np.where(m > 1e-32, np.floor(np.log2(np.maximum(m, 1e-32))) + 1, 0).astype(np.int64)
    scale = np.where(m > 1e-32, 256.0 / np.exp2(e), 0.0)
    mant = np.clip(np.floor(img * scale[..., None]), 0, 255).astype(np.uint8)
    rgbe = np.concatenate([mant, np.where(m > 1e-32, e + 128, 0).astype(np.uint8)[..., None]], axis=2)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
        for y in range(h):
            f.write(bytes([2, 2, w >> 8, w & 0xFF]))
            for c in range(4):
                row = rgbe[y, :, c]
                x = 0
                while x < w:
                    run = 1
                    while x + run < w and run < 127 and row[x + run] == row[x]:
                        run += 1
                    if run >= 3:
                        f.write(bytes([128 + run, int(row[x])]))
                        x += run
                    else:
                        n = 1
                        while x + n < w and n < 128 and not (x + n + 2 < w and row[x + n] == row[x + n + 1] == row[x + n + 2]):
                            n += 1
                        f.write(bytes([n]) + bytes(row[x:x + n].tolist()))
                        x += n


def write_inputs():
    os.makedirs(OUT, exist_ok=True)
    for name, text in (("box.obj", BOX_OBJ), ("tri.obj", TRI_OBJ), ("feature.scene", FEATURE_SCENE),
                       ("camera_matrix.scene", CAMERA_MATRIX_SCENE)):
        with open(os.path.join(OUT, name), "w") as f:
            f.write(text)
    write_png_files()
    y, x = np.mgrid[0:12, 0:24].astype(np.float64)
    sky = np.stack([0.2 + 0.05 * x, 0.3 + 0.02 * y, np.full_like(x, 0.9)], -1)
    sky[3:5, 5:8] = [40.0, 36.0, 30.0]  # a bright patch; long equal runs in the alpha channel
    write_hdr(os.path.join(OUT, "sky.hdr"), sky)


def write_expected():
    import refscene
    for name in ("feature", "camera_matrix"):
        r = refscene.load_scene_file(os.path.join(OUT, name + ".scene"))
        arrays = {k: v for k, v in r.items() if isinstance(v, np.ndarray)}
        for i, t in enumerate(r["textures"]):
            arrays[f"texture{i}"] = t
        arrays["num_textures"] = np.array(len(r["textures"]))
        arrays["top_level_index"] = np.array(r["top_level_index"])
        arrays["tex_array"] = np.array(r["tex_array"])
        arrays["has_env"] = np.array(r["has_env"])
        np.savez_compressed(os.path.join(OUT, f"expected_{name}.npz"), **arrays)
        print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    write_inputs()
    write_expected()
