"""Checkpoint directories (src/ref/reference_renderer.cpp:162-324) and OpenEXR output
(reference_renderer.cpp:673-705, src/lib/image.cpp:163-224) - CPU only.

The EXR files are read back by an independent reader written from the OpenEXR layout specification
(header attributes, offset table, scanline blocks, ZIP = zlib of the even/odd split byte-delta coded
block) and the HALF values compared with numpy's IEEE float16 conversion."""
import os
import struct
import zlib

import numpy as np
import pytest

import yar_amd
from yar_amd import film_dtype


def _tile(w, h, seed):
    rng = np.random.default_rng(seed)
    px = np.zeros((h, w), dtype=film_dtype())
    px["color_sum"] = rng.uniform(0, 5, (h, w, 3)).astype(np.float32)
    px["weight_sum"] = rng.integers(1, 9, (h, w)).astype(np.float32)
    return px


def test_checkpoint_new_directory_metadata(tmp_path):
    d = tmp_path / "ck" / "nested"
    ck = yar_amd.Checkpoint(str(d), "scene.pbrt", 16, 64)
    assert ck.finished_tiles() == [] and ck.previous_time == 0.0
    assert (d / "checkpoint").read_text() == "input_filename scene.pbrt\ntotal_tile_count 16\nsamples_per_pixer 64\n"


def test_checkpoint_tile_file_layout_and_resume(tmp_path):
    d = tmp_path / "ck"
    yar_amd.Checkpoint(str(d), "scene.pbrt", 300, 16).close()
    tiles = {3: ((63, 127, 129, 193), 1.5, 2.25), 17: ((0, 0, 65, 64), 0.25, 7.5), 250: ((10, 20, 12, 21), 3.0, 1.0)}
    data = {}
    for idx, (b, var, t) in tiles.items():
        px = _tile(b[2] - b[0], b[3] - b[1], idx)
        data[idx] = px
        yar_amd.checkpoint_write_tile(str(d), idx, b, px, var, t)
    # byte layout: f32 time, f64 variance, i32 x0 y0 x1 y1, Film_Pixels (16 B) - no padding
    raw = (d / "tile_0003").read_bytes()
    b, var, t = tiles[3]
    assert len(raw) == 4 + 8 + 16 + 16 * (b[2] - b[0]) * (b[3] - b[1])
    assert struct.unpack("<f", raw[:4])[0] == t and struct.unpack("<d", raw[4:12])[0] == var
    assert struct.unpack("<4i", raw[12:28]) == b
    assert raw[28:] == data[3].tobytes()
    assert sorted(os.listdir(d)) == ["checkpoint", "tile_0003", "tile_0017", "tile_0250"]   # temp files renamed
    ck = yar_amd.Checkpoint(str(d), "scene.pbrt", 300, 16)
    assert ck.finished_tiles() == [3, 17, 250]
    assert ck.previous_time == 7.5                                   # the latest session time stored
    for idx, (b, var, _) in tiles.items():
        rb, px, rv = ck.tile(idx)
        assert rb == b and rv == var and px.tobytes() == data[idx].tobytes()
    with pytest.raises(yar_amd.YarError):
        ck.tile(4)
    # a tile file is never overwritten
    with pytest.raises(yar_amd.YarError):
        yar_amd.checkpoint_write_tile(str(d), 17, tiles[17][0], data[17], 0.0, 0.0)
    # five digits for indices past 9999
    yar_amd.Checkpoint(str(tmp_path / "big"), "s", 20000, 1).close()
    yar_amd.checkpoint_write_tile(str(tmp_path / "big"), 12345, (0, 0, 1, 1), _tile(1, 1, 0), 0.0, 0.0)
    assert (tmp_path / "big" / "tile_12345").exists()


@pytest.mark.parametrize("field,args", [("input file", ("other.pbrt", 300, 16)), ("tile count", ("scene.pbrt", 299, 16)),
                                        ("spp", ("scene.pbrt", 300, 64))])
def test_checkpoint_rejects_changed_project(tmp_path, field, args):
    d = tmp_path / "ck"
    yar_amd.Checkpoint(str(d), "scene.pbrt", 300, 16).close()
    with pytest.raises(yar_amd.YarError) as e:
        yar_amd.Checkpoint(str(d), *args)
    assert e.value.code == 1 and "can not resume" in str(e.value)


def test_checkpoint_rejects_foreign_directory(tmp_path):
    (tmp_path / "x").mkdir()
    (tmp_path / "x" / "notes.txt").write_text("hello")
    with pytest.raises(yar_amd.YarError) as e:
        yar_amd.Checkpoint(str(tmp_path / "x"), "scene.pbrt", 4, 1)
    assert "not a checkpoint directory" in str(e.value)


def read_exr(path):
    """Minimal single-part scanline OpenEXR reader (HALF channels, NO or ZIP compression)."""
    b = open(path, "rb").read()
    assert struct.unpack("<I", b[:4])[0] == 20000630 and struct.unpack("<I", b[4:8])[0] == 2
    pos, attrs, order = 8, {}, []
    while b[pos] != 0:
        e = b.index(b"\0", pos)
        name = b[pos:e].decode()
        e2 = b.index(b"\0", e + 1)
        typ = b[e + 1:e2].decode()
        size = struct.unpack("<i", b[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + size])
        order.append(name)
        pos = e2 + 5 + size
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chl, chans = attrs["channels"][1], []
    p = 0
    while chl[p] != 0:
        e = chl.index(b"\0", p)
        ptype = struct.unpack("<i", chl[e + 1:e + 5])[0]
        chans.append((chl[p:e].decode(), ptype))
        p = e + 17
    comp = attrs["compression"][1][0]
    lines = {0: 1, 3: 16}[comp]
    nblocks = (h + lines - 1) // lines
    offsets = struct.unpack(f"<{nblocks}Q", b[pos:pos + 8 * nblocks])
    img = {c: np.zeros((h, w), np.float16) for c, _ in chans}
    for off in offsets:
        y, size = struct.unpack("<ii", b[off:off + 8])
        data = b[off + 8:off + 8 + size]
        nl = min(lines, h - y)
        raw_size = nl * w * 2 * len(chans)
        if comp == 3 and size < raw_size:
            t = bytearray(zlib.decompress(data))
            for i in range(1, len(t)):
                t[i] = (t[i - 1] + t[i] - 128) & 0xFF
            half = (len(t) + 1) // 2
            raw = bytearray(len(t))
            raw[0::2] = t[:half]
            raw[1::2] = t[half:]
            data = bytes(raw)
        arr = np.frombuffer(data, np.float16).reshape(nl, len(chans), w)
        for k, (c, _) in enumerate(chans):
            img[c][y:y + nl] = arr[:, k]
    return attrs, order, chans, img


def _attr_value(attrs, name):
    typ, v = attrs[name]
    return {"string": lambda: v.decode(), "int": lambda: struct.unpack("<i", v)[0],
            "float": lambda: struct.unpack("<f", v)[0]}[typ]()


@pytest.mark.parametrize("compress", [False, True])
def test_exr_half_channels_and_attributes(tmp_path, compress):
    rng = np.random.default_rng(3)
    h, w = 37, 53
    img = rng.uniform(0, 4, (h, w, 3)).astype(np.float32)
    img[0, :4, 0] = [0.0, 65504.0, 70000.0, 1e-7]                   # max half, overflow, subnormal
    img[1, :3, 1] = [6.1e-5, 2.0 ** -25, 3 * 2.0 ** -26]              # around the smallest normal / half ulp ties
    img[2, :2, 2] = [np.inf, 1.0 + 2.0 ** -11]                        # infinity, a round-to-even tie
    path = str(tmp_path / "img.exr")
    yar_amd.write_exr(path, img, input_file="scenes/x.pbrt", spp=64, variance=0.125, load_time=1.5,
                      render_time=2.5, compress=compress)
    attrs, order, chans, out = read_exr(path)
    assert chans == [("B", 1), ("G", 1), ("R", 1)]
    assert attrs["compression"][1][0] == (3 if compress else 0)
    with np.errstate(over="ignore"):
        want = img.astype(np.float16)                                # IEEE round to nearest even
    for k, c in enumerate("RGB"):
        assert np.array_equal(out[c].view(np.uint16), want[..., k].view(np.uint16)), c
    assert _attr_value(attrs, "yar_input_file") == "scenes/x.pbrt"
    assert _attr_value(attrs, "yar_spp") == 64 and _attr_value(attrs, "yar_variance") == 0.125
    assert _attr_value(attrs, "yar_build_version") == "0.0" and _attr_value(attrs, "yar_render_device") == "gpu"
    assert "yar_load_time" not in attrs and "yar_render_time" not in attrs   # varying attributes off by default
    assert order[:8] == ["channels", "compression", "dataWindow", "displayWindow", "lineOrder", "pixelAspectRatio",
                         "screenWindowCenter", "screenWindowWidth"]


def test_exr_varying_attributes_and_dump(tmp_path):
    path = str(tmp_path / "r.exr")
    yar_amd.write_exr(path, np.ones((4, 4, 3), np.float32), input_file="s.pbrt", spp=4, variance=0.5, load_time=1.25,
                      render_time=3.5, varying_attributes=True, dump_attributes=True)
    attrs, _, _, _ = read_exr(path)
    assert _attr_value(attrs, "yar_load_time") == 1.25 and _attr_value(attrs, "yar_render_time") == 3.5
    txt = (tmp_path / "r.txt").read_text().splitlines()
    assert txt[0] == "yar_build_version 0.0" and "yar_spp 4" in txt and "yar_render_time 3.500000" in txt
