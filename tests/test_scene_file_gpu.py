"""Scene files in the reference schema (two-document YAML + OBJ) load into the same device scene
as the arrays they were written from, and the command-line renderer writes an EXR."""
import os
import subprocess

import numpy as np
import pytest

import swr
from swr import scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_yaml_scene_renders_like_the_arrays(ctx, tmp_path):
    desc = scenes.box_room(width=40, height=32, base=2, sphere_level=2)
    path = scenes.write_scene(desc, str(tmp_path), "room")
    sc_file, rd_file = swr.load_yaml(ctx, path)
    sc_arr = swr.Scene(ctx, desc)
    assert sc_file.num_triangles == sc_arr.num_triangles
    assert sc_file.num_geometries == sc_arr.num_geometries
    rd = swr.render_desc(desc["render"])
    for f in ("width", "height", "base", "sampler", "filter", "min_depth", "max_depth", "bin_exponent"):
        assert getattr(rd_file, f) == getattr(rd, f), f
    assert rd_file.threshold == np.float32(rd.threshold)
    np.testing.assert_allclose(list(rd_file.camera.translation), list(rd.camera.translation), rtol=1e-7)
    rd_file.seed = rd.seed
    p_file, _, _ = sc_file.render(rd_file)
    p_arr, _, _ = sc_arr.render(rd)
    np.testing.assert_array_equal(p_file, p_arr)


def test_textured_yaml_scene_renders_like_the_arrays(ctx, tmp_path):
    """Standard, Layered and Constant textures through the scene file (images as PFM, which
    holds float32 exactly): the same image as the scene built from arrays, bit for bit."""
    desc = scenes.textured_room(width=40, height=32, base=2, sphere_level=2)
    path = scenes.write_scene(desc, str(tmp_path), "tex")
    sc_file, rd_file = swr.load_yaml(ctx, path)
    sc_arr = swr.Scene(ctx, desc)
    rd = swr.render_desc(desc["render"])
    rd_file.seed = rd.seed
    p_file, _, _ = sc_file.render(rd_file)
    p_arr, _, _ = sc_arr.render(rd)
    np.testing.assert_array_equal(p_file, p_arr)
    sc_file.close()
    sc_arr.close()


def _write_png_rgb8(path, rgb8):
    """An 8-bit RGB PNG (ISO/IEC 15948: IHDR, one IDAT of unfiltered rows, IEND)."""
    import struct
    import zlib
    h, w, _ = rgb8.shape
    raw = b"".join(b"\0" + rgb8[y].astype(np.uint8).tobytes() for y in range(h))

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
                + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def test_png_textures_render_like_pfm(ctx, tmp_path):
    """The textured room with its images quantised to 8 bits, once with the images as PFM and once
    as PNG (a StandardTexture string naming a .png, as reference scene files do): the same image,
    bit for bit (the PNG reader's sample / 255 is the value the PFM holds)."""
    desc = scenes.textured_room(width=40, height=32, base=2, sphere_level=2)
    q = [np.clip(np.rint(np.asarray(im, np.float32) * 255), 0, 255).astype(np.uint8) for im in desc["images"]]
    desc["images"] = [x.astype(np.float32) / np.float32(255) for x in q]
    path = scenes.write_scene(desc, str(tmp_path), "tex")
    text = open(path).read()
    for k, x in enumerate(q):
        _write_png_rgb8(str(tmp_path / ("tex_image%d.png" % k)), x)
        assert ("tex_image%d.pfm" % k) in text
        text = text.replace("tex_image%d.pfm" % k, "tex_image%d.png" % k)
    png_path = str(tmp_path / "tex_png.yaml")
    with open(png_path, "w") as f:
        f.write(text)
    sc_pfm, rd_pfm = swr.load_yaml(ctx, path)
    sc_png, rd_png = swr.load_yaml(ctx, png_path)
    rd_png.seed = rd_pfm.seed
    p_pfm, _, _ = sc_pfm.render(rd_pfm)
    p_png, _, _ = sc_png.render(rd_png)
    assert np.isfinite(p_png).all() and p_png.max() > 0
    np.testing.assert_array_equal(p_png, p_pfm)
    sc_pfm.close()
    sc_png.close()


def test_yaml_texture_errors(ctx, tmp_path):
    (tmp_path / "t.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n")
    obj = ("object: {type: Polygon, filename: t.obj, translation: [0, 0, 0], rotation: [0, 0, 0], "
           "scale: [1, 1, 1], shader: 0}\n")
    cases = [("{type: Standard, string: missing.pfm}", swr.SWR_E_IO),
             ("{type: Checker, colour: [1, 1, 1]}", swr.SWR_E_PARSE),
             ("{type: Layered, upper: {type: Constant, colour: [1, 1, 1]}, lower: {type: Constant, "
              "colour: [0, 0, 0]}}", swr.SWR_E_PARSE)]
    for k, (tex, code) in enumerate(cases):
        (tmp_path / ("s%d.yaml" % k)).write_text(
            "{}\n---\nshader: {type: Lambert, texture: %s, reflectance: 1}\n%s" % (tex, obj))
        with pytest.raises(swr.SwrError) as e:
            swr.load_yaml(ctx, str(tmp_path / ("s%d.yaml" % k)))
        assert e.value.code == code, tex


def test_yaml_defaults_and_strict_decoders(ctx, tmp_path):
    (tmp_path / "t.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "ok.yaml").write_text(
        "{}\n---\nshader: {type: Lambert, texture: {type: Constant, colour: [1, 1, 1]}, reflectance: 1}\n"
        "object: {type: Polygon, filename: t.obj, translation: [0, 0, 0], rotation: [0, 0, 0], scale: [1, 1, 1], "
        "shader: 0}\n")
    sc, rd = swr.load_yaml(ctx, str(tmp_path / "ok.yaml"))
    # defaults of Pathtracer.cpp:45-161
    assert (rd.width, rd.height, rd.base, rd.min_depth, rd.max_depth) == (700, 500, 8, 2, 100)
    assert rd.camera.type == swr.CAMERA_THINLENS and rd.camera.focal_length == 55.0
    assert sc.num_triangles == 1
    (tmp_path / "bad.yaml").write_text("image: {width: 10, height: 10}\n---\n{}\n")  # 2 of 3 keys
    with pytest.raises(swr.SwrError) as e:
        swr.load_yaml(ctx, str(tmp_path / "bad.yaml"))
    assert e.value.code == swr.SWR_E_PARSE
    with pytest.raises(swr.SwrError) as e:
        swr.load_yaml(ctx, str(tmp_path / "missing.yaml"))
    assert e.value.code == swr.SWR_E_IO


def test_cli_writes_exr(tmp_path):
    exe = os.path.join(ROOT, "msc-project_amd", "bin", "swr")
    if not os.path.exists(exe):
        pytest.skip("command-line renderer not built")
    desc = scenes.box_room(width=32, height=24, base=2, sphere_level=1)
    path = scenes.write_scene(desc, str(tmp_path), "room")
    out = tmp_path / "o.exr"
    r = subprocess.run([exe, "-o", str(out), path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = out.read_bytes()
    assert data[:4] == bytes([0x76, 0x2F, 0x31, 0x01])


def test_cli_progressive_preview_is_the_viewer_image(tmp_path):
    """Without -o the command line runs the viewer's progressive loop headless: -n iterations,
    each adding one process() to the framebuffer, and -p writes the viewer's 8-bit image,
    min(1, sum / iterations) * 255 truncated (main.cpp:163-168), which must match the sum of the
    same iterations rendered through the Python facade."""
    exe = os.path.join(ROOT, "msc-project_amd", "bin", "swr")
    if not os.path.exists(exe):
        pytest.skip("command-line renderer not built")
    desc = scenes.box_room(width=20, height=12, base=2, sphere_level=1)
    path = scenes.write_scene(desc, str(tmp_path), "room")
    out = tmp_path / "preview.ppm"
    r = subprocess.run([exe, "-n", "3", "-p", str(out), path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("is complete") == 3
    data = out.read_bytes()
    header = b"P6\n20 12\n255\n"
    assert data[:len(header)] == header
    got = np.frombuffer(data[len(header):], np.uint8).reshape(12, 20, 3)
    pt = swr.Pathtracer(path)
    for _ in range(3):
        pt.process()
    total, _, _ = pt.image()
    want = (np.minimum(np.float32(1.0), total / np.float32(3.0)) * np.float32(255.0)).astype(np.uint8)
    np.testing.assert_array_equal(got, want)
    bad = subprocess.run([exe, "-n", "0", path], capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "positive integer" in bad.stderr
