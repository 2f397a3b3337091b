"""Scene ingestion: OBJ meshes and the reference's XML scene description
(src/io/objfilereader.cpp:117-217, src/io/customfilereader.cpp:47-340).  The
reference ships no scene files, so the fixtures are written here in the
formats its readers parse; expectations come from the file text."""
import numpy as np
import pytest

from renderbox2_amd import _lib as L

CUBE_ROOM_OBJ = """# two objects; faces use every corner syntax of the format
o floor
v -1 0 -1
v  1 0 -1
v  1 0  1
v -1 0  1
vt 0 0
vt 1 0
vn 0 1 0
f 1/1/1 2/2/1 3/1/1
f 1//1 3//1 4//1
o lamp
v -0.25 1.9 -0.25   # a trailing comment
v  0.25 1.9 -0.25
v  0.25 1.9  0.25
v -0.25 1.9  0.25
f 5 8 7 6
o mirrorwall
v 1 0 -1
v 1 2 -1
v 1 2 1
f -3 -2 -1
"""

SCENE_XML = """<?xml version="1.0"?>
<!-- cornell-style test scene -->
<scene>
  <files>
    <count>1</count>
    <file>room.obj</file>
  </files>
  <camera>
    <position>{0, 1, -3}</position>
    <lookat>{0,1,0}</lookat>
    <up>0 1 0</up>
    <fov>45</fov>
  </camera>
  <material>
    <name>white</name>
    <bsdf><layers>1</layers>
      <bxdf><type>lambertian</type><reflectance>{0.7, 0.6, 0.5}</reflectance></bxdf>
    </bsdf>
  </material>
  <material>
    <name>silver</name>
    <bsdf><layers>1</layers>
      <bxdf><type>mirror</type><reflectance>{0.9,0.9,0.9}</reflectance></bxdf>
    </bsdf>
  </material>
  <object><name>floor</name><material>white</material></object>
  <object><name>mirrorwall</name><material>silver</material></object>
  <lightobjects>
    <light><name>lamp</name><type>diffuse</type><Le>{10, 9, 8}</Le></light>
  </lightobjects>
</scene>
"""


def _write(tmp_path, obj=CUBE_ROOM_OBJ, xml=SCENE_XML):
    (tmp_path / "room.obj").write_text(obj)
    (tmp_path / "room.xml").write_text(xml)
    return tmp_path / "room.xml"


def test_obj_reader_keeps_file_order_and_global_indices(rbm, tmp_path):
    _write(tmp_path)
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((0.5, 0.5, 0.5))
    s.load_obj(tmp_path / "room.obj", m)
    v, idx, tm, lt = s.geometry()
    assert v.shape == (11, 3) and np.allclose(v[4], [-0.25, 1.9, -0.25]) and np.allclose(v[10], [1, 2, 1])
    # floor: 2 triangles; lamp quad fanned (5 8 7 6 -> 5 8 7, 5 7 6); wall: -3 -2 -1 = 9 10 11
    expect = [[0, 1, 2], [0, 2, 3], [4, 7, 6], [4, 6, 5], [8, 9, 10]]
    assert idx.tolist() == expect
    assert (tm == m).all() and len(lt) == 0


def test_xml_scene_materials_lights_and_camera(rbm, tmp_path):
    path = _write(tmp_path)
    s = rbm.Scene()
    cam = s.load_xml(path, 320, 240)
    v, idx, tm, lt = s.geometry()
    mats = [m[:2] for m in s.materials()]
    # materials in file order, then one emitter per light
    assert mats == [(L.RB_MAT_LAMBERT, pytest.approx((0.7, 0.6, 0.5))), (L.RB_MAT_MIRROR, pytest.approx((0.9,) * 3)),
                    (L.RB_MAT_EMITTER, pytest.approx((10.0, 9.0, 8.0)))]
    assert tm.tolist() == [0, 0, 2, 2, 1]
    assert lt.tolist() == [2, 3]
    c = cam.c
    assert list(c.position) == [0, 1, -3] and list(c.lookat) == [0, 1, 0] and list(c.up) == [0, 1, 0]
    assert c.hfov_deg == 45.0 and (c.width, c.height) == (320, 240)


def test_xml_scene_equals_the_same_scene_built_by_hand(rbm, tmp_path):
    path = _write(tmp_path)
    a = rbm.Scene()
    a.load_xml(path)
    b = rbm.Scene()
    w = b.add_diffuse_bsdf((0.7, 0.6, 0.5))
    mi = b.add_mirror_bsdf((0.9, 0.9, 0.9))
    e = b.add_diffuse_light_bsdf((10.0, 9.0, 8.0))
    va, ia, _, _ = a.geometry()
    b.add_mesh(va[0:4], [[0, 1, 2], [0, 2, 3]], w)
    b.add_mesh(va[4:8], [[0, 3, 2], [0, 2, 1]], e)
    b.add_mesh(va[8:11], [[0, 1, 2]], mi)
    for x, y in zip(a.geometry(), b.geometry()):
        assert np.array_equal(x, y)
    ba = rbm.BVH(a)
    bb = rbm.BVH(b)
    for x, y in zip(ba.export(L.RB_LAYOUT_MI355X), bb.export(L.RB_LAYOUT_MI355X)):
        assert np.array_equal(x, y)


def test_scene_file_errors(rbm, tmp_path):
    s = rbm.Scene()
    s.add_diffuse_bsdf((1, 1, 1))
    with pytest.raises(rbm.RbError):
        s.load_obj(tmp_path / "missing.obj", 0)
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(rbm.RbError):  # vertex 3 does not exist
        s.load_obj(tmp_path / "bad.obj", 0)
    layered = SCENE_XML.replace("<bxdf><type>mirror</type><reflectance>{0.9,0.9,0.9}</reflectance></bxdf>",
                                "<bxdf><type>mirror</type><reflectance>{0.9,0.9,0.9}</reflectance></bxdf>"
                                "<bxdf><type>lambertian</type><reflectance>{0.1,0.1,0.1}</reflectance></bxdf>")
    with pytest.raises(rbm.RbError) as e:  # <layers>1</layers> with two <bxdf>
        rbm.Scene().load_xml(_write(tmp_path, xml=layered))
    assert e.value.status == L.RB_ERR_INVALID
    five = SCENE_XML.replace("<bxdf><type>mirror</type><reflectance>{0.9,0.9,0.9}</reflectance></bxdf>",
                             "<bxdf><type>mirror</type><reflectance>{0.9,0.9,0.9}</reflectance></bxdf>" * 5)
    with pytest.raises(rbm.RbError) as e:  # more than four layers
        rbm.Scene().load_xml(_write(tmp_path, xml=five.replace("<name>silver</name>\n    <bsdf><layers>1</layers>",
                                                               "<name>silver</name>\n    <bsdf><layers>5</layers>")))
    assert e.value.status == L.RB_ERR_INVALID and "1..4" in str(e.value)
    unknown = SCENE_XML.replace("<type>mirror</type>", "<type>velvet</type>")
    with pytest.raises(rbm.RbError) as e:
        rbm.Scene().load_xml(_write(tmp_path, xml=unknown))
    assert e.value.status == L.RB_ERR_UNSUPPORTED
    glass_no_eta = SCENE_XML.replace("<type>mirror</type>", "<type>glass</type>")
    with pytest.raises(rbm.RbError) as e:
        rbm.Scene().load_xml(_write(tmp_path, xml=glass_no_eta))
    assert e.value.status == L.RB_ERR_INVALID
    with pytest.raises(rbm.RbError) as e:
        rbm.Scene().load_xml(_write(tmp_path, xml=SCENE_XML.replace("</scene>", "")))
    assert e.value.status == L.RB_ERR_INVALID
    unknown = SCENE_XML.replace("<material>silver</material>", "<material>gold</material>")
    with pytest.raises(rbm.RbError):
        rbm.Scene().load_xml(_write(tmp_path, xml=unknown))


# ------------------------------------------------------------ image output --

def _img():
    rng = np.random.default_rng(5)
    img = rng.uniform(-0.2, 1.3, (7, 9, 3)).astype(np.float32)
    img[0, 0] = [0.0, -0.0, 1.0]
    img[0, 1] = [65504.0, 70000.0, 1e-7]    # max half, overflow, subnormal
    img[0, 2] = [6e-8, 2.9802322e-08, 1e-9]  # smallest subnormal, tie at half of it, underflow
    img[1, 0] = [1.0009765625 + 2 ** -11, 1.0 + 2 ** -11, 0.5]  # ties: to even
    return img


def test_pfm_round_trip_is_exact(rbm, tmp_path):
    img = _img()
    rbm.write_image(tmp_path / "a.pfm", img)
    raw = (tmp_path / "a.pfm").read_bytes()
    head, rest = raw.split(b"\n", 3)[:3], raw.split(b"\n", 3)[3]
    assert head[0] == b"PF" and head[1] == b"9 7" and float(head[2]) < 0
    back = np.frombuffer(rest, "<f4").reshape(7, 9, 3)[::-1]
    assert np.array_equal(back.view(np.uint32), img.view(np.uint32))


def _to_byte(img):
    c = np.clip(img, 0.0, 1.0).astype(np.float32)
    return (np.power(c, np.float32(1.0 / 2.2)) * np.float32(255) + 0.5).astype(np.int64)


def test_png_and_ppm_are_valid_and_agree(rbm, tmp_path):
    import zlib
    img = _img()
    rbm.write_image(tmp_path / "a.ppm", img)
    rbm.write_image(tmp_path / "a.png", img)
    ppm = (tmp_path / "a.ppm").read_bytes()
    assert ppm.startswith(b"P6\n9 7\n255\n")
    pix = np.frombuffer(ppm[len(b"P6\n9 7\n255\n"):], np.uint8).reshape(7, 9, 3)
    assert np.abs(pix.astype(np.int64) - _to_byte(img)).max() <= 1  # libm powf vs numpy float32 power
    png = (tmp_path / "a.png").read_bytes()
    assert png[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(png):
        n = int.from_bytes(png[pos:pos + 4], "big")
        typ, data = png[pos + 4:pos + 8], png[pos + 8:pos + 8 + n]
        crc = int.from_bytes(png[pos + 8 + n:pos + 12 + n], "big")
        assert zlib.crc32(typ + data) == crc
        chunks[typ] = data
        pos += 12 + n
    assert set(chunks) == {b"IHDR", b"IDAT", b"IEND"}
    assert chunks[b"IHDR"] == (9).to_bytes(4, "big") + (7).to_bytes(4, "big") + bytes([8, 2, 0, 0, 0])
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(7, 1 + 27)
    assert (raw[:, 0] == 0).all()
    assert np.array_equal(raw[:, 1:].reshape(7, 9, 3), pix)


def test_exr_holds_the_half_values(rbm, tmp_path):
    img = _img()
    rbm.write_image(tmp_path / "a.exr", img)
    b = (tmp_path / "a.exr").read_bytes()
    assert int.from_bytes(b[:4], "little") == 20000630 and int.from_bytes(b[4:8], "little") == 2
    pos, attrs = 8, {}
    while b[pos] != 0:
        name_end = b.index(b"\0", pos)
        type_end = b.index(b"\0", name_end + 1)
        size = int.from_bytes(b[type_end + 1:type_end + 5], "little")
        attrs[b[pos:name_end].decode()] = (b[name_end + 1:type_end].decode(), b[type_end + 5:type_end + 5 + size])
        pos = type_end + 5 + size
    pos += 1
    assert attrs["compression"][1] == b"\0"
    assert np.frombuffer(attrs["dataWindow"][1], "<i4").tolist() == [0, 0, 8, 6]
    ch = attrs["channels"][1]
    assert ch.startswith(b"B\0") and b"G\0" in ch and b"R\0" in ch
    offsets = np.frombuffer(b[pos:pos + 8 * 7], "<u8")
    with np.errstate(over="ignore"):
        half = img.astype(np.float16)  # round to nearest even; 70000 -> inf
    for y in range(7):
        o = int(offsets[y])
        assert int.from_bytes(b[o:o + 4], "little") == y and int.from_bytes(b[o + 4:o + 8], "little") == 9 * 3 * 2
        line = np.frombuffer(b[o + 8:o + 8 + 54], "<u2").reshape(3, 9)  # B, G, R
        for c, k in ((0, 2), (1, 1), (2, 0)):
            assert np.array_equal(line[c], half[y, :, k].view(np.uint16)), (y, k)


@pytest.mark.parametrize("shape", [(29, 37), (8, 8), (1, 1), (17, 3)])
def test_jpeg_decodes_to_the_8bit_image(rbm, tmp_path, shape):
    # a baseline JPEG that an independent decoder (Pillow) reads back: at the
    # reference's quality 100 every quantiser step is 1, so the decoded pixels
    # are the 8-bit image up to the colour transform's and the DCT's rounding
    Image = pytest.importorskip("PIL.Image")
    h, w = shape
    y, x = np.mgrid[0:h, 0:w]
    rng = np.random.default_rng(h * 100 + w)
    img = np.stack([x / max(w - 1, 1), y / max(h - 1, 1), 0.5 + 0.5 * np.sin(x * 0.7 + y * 0.3)], -1)
    img = (img + rng.uniform(-0.1, 0.1, img.shape)).astype(np.float32)
    img[0, 0] = [1.5, -0.5, 0.25]  # clamped like the 8-bit writers
    rbm.write_image(tmp_path / "a.jpg", img)
    data = (tmp_path / "a.jpg").read_bytes()
    assert data[:4] == b"\xff\xd8\xff\xe0" and data[-2:] == b"\xff\xd9"
    with Image.open(tmp_path / "a.jpg") as im:
        assert im.format == "JPEG" and im.mode == "RGB" and im.size == (w, h)
        got = np.asarray(im, np.int64)
    err = np.abs(got - _to_byte(img))
    assert err.max() <= 3 and err.mean() < 1.0, (err.max(), err.mean())


def test_jpeg_of_a_flat_image_is_exact(rbm, tmp_path):
    Image = pytest.importorskip("PIL.Image")
    img = np.full((16, 24, 3), 0.5, np.float32)
    rbm.write_image(tmp_path / "flat.jpeg", img)
    with Image.open(tmp_path / "flat.jpeg") as im:
        got = np.asarray(im, np.int64)
    assert np.abs(got - _to_byte(img)).max() <= 1


def test_image_write_rejects_unknown_suffix(rbm, tmp_path):
    with pytest.raises(rbm.RbError):
        rbm.write_image(tmp_path / "a.tiff", _img())


def test_xml_reads_glass_orennayar_and_microfacet(rbm, tmp_path):
    xml = SCENE_XML.replace("""  <object><name>floor</name>""", """  <material>
    <name>pane</name>
    <bsdf><layers>1</layers><bxdf><type>glass</type><reflectance>{1,1,1}</reflectance>
      <etai>1.0</etai><etat>1.5</etat></bxdf></bsdf>
  </material>
  <material>
    <name>clay</name>
    <bsdf><layers>1</layers><bxdf><type>orennayar</type><reflectance>{0.5,0.4,0.3}</reflectance>
      <sigma>20</sigma></bxdf></bsdf>
  </material>
  <material>
    <name>gold</name>
    <bsdf><layers>1</layers><bxdf><type>microfacet</type><reflectance>{1,1,1}</reflectance>
      <distribution>blinn</distribution><exponent>40</exponent><fresnel_type>fresnel_conductor</fresnel_type>
      <eta>{0.2, 0.4, 1.4}</eta><k>{3.6, 2.4, 1.8}</k></bxdf></bsdf>
  </material>
  <object><name>floor</name>""").replace("<object><name>floor</name><material>white</material>",
                                          "<object><name>floor</name><material>clay</material>")
    s = rbm.Scene()
    s.load_xml(_write(tmp_path, xml=xml))
    mats = s.materials()
    kinds = [m[0] for m in mats]
    assert kinds == [L.RB_MAT_LAMBERT, L.RB_MAT_MIRROR, L.RB_MAT_GLASS, L.RB_MAT_ORENNAYAR, L.RB_MAT_MICROFACET,
                     L.RB_MAT_EMITTER]
    assert mats[2][2][:2] == (1.0, 1.5)
    sig = np.float32(20.0) * np.float32(np.pi / 180)
    s2 = np.float32(sig * sig)
    assert mats[3][2][0] == 20.0
    assert mats[3][2][1] == pytest.approx(1 - s2 / (2 * (s2 + 0.33)), rel=1e-6)  # A
    assert mats[3][2][2] == pytest.approx(0.45 * s2 / (s2 + 0.09), rel=1e-6)     # B
    assert mats[4][2][:2] == (40.0, L.RB_FRESNEL_CONDUCTOR)
    assert mats[4][2][2:8] == pytest.approx((0.2, 0.4, 1.4, 3.6, 2.4, 1.8))
    _, _, tm, _ = s.geometry()
    assert tm.tolist()[:2] == [3, 3]  # the floor got the Oren-Nayar material
