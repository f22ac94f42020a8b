"""Host-side behaviour of the wr:: API that needs no GPU: error messages of
libWR's host-forcible failures (scene.cpp:171-173,249-251,356-358) and the
PNG reader on the reference's own texture fixture."""
from pathlib import Path

import pytest

import offlinerenderer_amd as wr

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_invalid_material_id():
    s = wr.Scene()
    s.addLambertianMaterial((0.5, 0.5, 0.5))
    with pytest.raises(RuntimeError, match="Invalid material ID"):
        s.setMaterialNormalMap(5, 1)


def test_empty_scene_finalize():
    s = wr.Scene()
    with pytest.raises(RuntimeError, match="Scene has no geometry"):
        s.finalize()


def test_missing_texture_message(tmp_path):
    s = wr.Scene()
    with pytest.raises(RuntimeError, match="Failed to load texture: .*nope.png"):
        wr.Texture2D.create(s, str(tmp_path / "nope.png"))
    bad = tmp_path / "bad.png"
    bad.write_bytes(b"not a png")
    with pytest.raises(RuntimeError, match="Failed to load texture"):
        wr.Texture2D.create(s, str(bad))


def test_reference_fixture_png_loads():
    s = wr.Scene()
    a = wr.Texture2D.create(s, str(GOLDEN / "checkerboard_line.png"))
    b = wr.Texture2D.create(s, str(GOLDEN / "checkerboard_line.png"))
    assert (a, b) == (1, 2)  # 1-based ids (scene.cpp:261)
    wr.Texture2D.destroy(s, a)


def test_material_ids_are_sequential():
    s = wr.Scene()
    ids = [s.addLambertianMaterial((0.1, 0.2, 0.3)), s.addEmissiveMaterial((1, 1, 1)),
           s.addGlassMaterial((1, 1, 1), 1.5), s.addGGXReflectionMaterial((1, 1, 1), 0.3),
           s.addGGXTransmissionMaterial((1, 1, 1), 0.2, 1.5)]
    assert ids == [0, 1, 2, 3, 4]


def test_missing_hdr_message(tmp_path):
    s = wr.Scene()
    with pytest.raises(RuntimeError, match="Failed to load HDR environment map"):
        s.setEnvironmentMap(str(tmp_path / "none.hdr"))
