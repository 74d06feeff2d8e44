"""Host-side logic of the product that runs without a GPU: the C-ABI
library loads and exports every declared symbol, the libconfig reader keeps
the reference's keys and error behaviour, status codes and messages cross the
boundary, and the seeded synthetic source is deterministic."""
from __future__ import annotations

import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(risa_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(risa):
    from risa_amd import _lib

    lib = _lib.load()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding binds exactly what the headers declare
    assert set(_lib.exported_symbols()) == names


def test_abi_version_and_geom_layout(risa):
    from risa_amd import _lib

    lib = _lib.load()
    assert lib.risa_abi_version() >= 1
    assert lib.risa_geom_size() == C.sizeof(_lib.Geom)


def test_config_c0_keys(risa):
    """Keys and derived values as the reference reads them (Attenuation.cu:387-409,
    Fan2Para.cu:595-619, Filter.cu:196-215, Backprojection.cu:214-230, Masking.cu:149-151)."""
    g = risa.geom_from_config(risa.config_path("c0"))
    assert g.numberOfFanDetectors == 432
    assert g.numberOfFanProjections == 500          # samplingRate * 1e6 / scanRate
    assert g.numberOfPlanes == 1
    assert g.numberOfParallelProjections == 500 and g.numberOfParallelDetectors == 256
    assert g.numberOfPixels == 256
    assert g.filterType == 0 and g.cutoffFraction == 1.0
    assert g.sourceOffset == 23.2 and g.detectorDiameter == 216.0 and g.imageWidth == 190.0
    assert g.sourceDiameter[0] == 365.0 and g.deltaX[0] == 815.0 and g.deltaZ[0] == 1417.0
    assert g.sourceAngle[0] == 240.0
    assert (g.xa, g.xb, g.xc, g.xd, g.xe, g.xf) == (43, 285, 175, 362, 0, 431)
    assert g.threshMin == 0.67 and g.threshMax == 1.5
    assert g.fusedMasking == 1 and g.normalization == 0


def test_config_c1_keys(risa):
    g = risa.geom_from_config(risa.config_path("c1"))
    assert g.numberOfFanDetectors == 768 and g.numberOfFanProjections == 1000
    assert g.numberOfPlanes == 2
    assert g.numberOfParallelProjections == 1000 and g.numberOfParallelDetectors == 512
    assert g.numberOfPixels == 512 and g.filterType == 1
    assert g.sourceDiameter[1] == 370.0 and g.deltaZ[1] == 1430.0


def write_cfg(tmp_path, text, name="t.cfg"):
    p = tmp_path / name
    p.write_text(text)
    return p


def c0_text():
    return (ROOT / "risa_amd" / "configs" / "c0.cfg").read_text()


def test_config_errors(risa, tmp_path):
    from risa_amd import _lib

    with pytest.raises(_lib.RisaError, match="RISA_ERR_CONFIG"):
        risa.geom_from_config(tmp_path / "missing.cfg")
    # a required key removed
    text = re.sub(r"numberOfParallelDetectors\s*=\s*256\n", "", c0_text())
    with pytest.raises(_lib.RisaError, match="numberOfParallelDetectors"):
        risa.geom_from_config(write_cfg(tmp_path, text))
    # a syntax error names the line
    with pytest.raises(_lib.RisaError, match="line"):
        risa.geom_from_config(write_cfg(tmp_path, c0_text() + "\nbroken = = 3\n", "b.cfg"))
    # an indexed key with too few elements for the planes
    text = c0_text().replace("numberOfPlanes = 1", "numberOfPlanes = 2")
    with pytest.raises(_lib.RisaError):
        risa.geom_from_config(write_cfg(tmp_path, text, "c.cfg"))


def test_config_unknown_names_fall_back_like_the_reference(risa, tmp_path):
    """Filter.cu:208-214: an unknown filterType means ramp; Backprojection.cu:
    224-228: an unknown interpolationType means linear."""
    text = c0_text().replace('filterType = "ramp"', 'filterType = "boxcar"')
    text = text.replace('interpolationType = "linear"', 'interpolationType = "cubic"')
    g = risa.geom_from_config(write_cfg(tmp_path, text))
    assert g.filterType == 0 and g.interpolationType == 0
    text = c0_text().replace('filterType = "ramp"', 'filterType = "hanning"')
    text = text.replace('interpolationType = "linear"', 'interpolationType = "nearestNeighbour"')
    g = risa.geom_from_config(write_cfg(tmp_path, text, "n.cfg"))
    assert g.filterType == 4 and g.interpolationType == 1


def test_config_accepts_reference_syntax(risa, tmp_path):
    """libconfig syntax as the reference's config.cfg uses it: `;` terminators,
    comments of three kinds, lists in [] and (), and unused keys (Q14)."""
    text = c0_text()
    text = text.replace("numberOfPixels = 256", "numberOfPixels = 256; /* block */ # hash")
    text = text.replace("sourceDiameter = [ 365.0 ]", "sourceDiameter = ( 365.0 );")
    text += '\ndetectorIPs = ["10.0.0.1", "10.0.0.2"];\nautoInterpolation = true;\nRofexVersion = 2\n'
    g = risa.geom_from_config(write_cfg(tmp_path, text))
    assert g.numberOfPixels == 256 and g.sourceDiameter[0] == 365.0


def test_status_codes_cross_the_boundary(risa):
    from risa_amd import _lib

    lib = _lib.load()
    rc = lib.risa_filter_taps(None, None)
    assert rc == 1                     # RISA_ERR_ARG
    assert lib.risa_last_error()


def test_reorder_and_layout_argument_errors(risa):
    """The reordering's argument checks run before any device work: frames
    narrower than 27 modules of 16 detectors are refused (the reference's
    table indexes past them), as are in-place calls and unknown layouts."""
    from risa_amd import _lib

    lib = _lib.load()
    dummy = C.c_void_p(4096)
    assert lib.risa_hip_reorder_run(None, 1, 10, 432, dummy, None) == 1          # RISA_ERR_ARG
    assert lib.risa_hip_reorder_run(dummy, 1, 10, 432, dummy, None) == 1         # in place
    assert lib.risa_hip_reorder_run(dummy, 1, 10, 400, C.c_void_p(8192), None) == 4   # RISA_ERR_UNSUPPORTED
    assert b"432" in lib.risa_last_error()
    assert lib.risa_hip_reorder_run(dummy, 0, 10, 432, C.c_void_p(8192), None) == 0   # nothing to do
    assert lib.risa_hip_attenuation_set_input_layout(None, 1) == 1
    assert lib.risa_hip_chain_set_input_layout(None, 1) == 1


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_tiff_writer_round_trip(risa, tmp_path, fmt):
    """risa_tiff_write_f32 (OfflineSaver.cu:135-178): BigTIFF sequences,
    big- and little-endian classic TIFF, read back by an independent reader
    (Pillow) page for page and bit for bit, with the reference's tags."""
    from PIL import Image, ImageSequence
    from risa_amd import _lib

    lib = _lib.load()
    imgs = np.random.default_rng(fmt).standard_normal((3, 24, 40)).astype(np.float32)
    imgs[0, 0, :3] = [np.inf, -0.0, 1e-40]     # special values pass through unchanged
    p = tmp_path / f"seq{fmt}.tif"
    _lib.check(lib.risa_tiff_write_f32(str(p).encode(), imgs.ctypes.data, 3, 40, 24, fmt))
    head = p.read_bytes()[:4]
    assert head == ((b"II+\x00") if fmt == 0 else (b"MM\x00*") if fmt == 1 else (b"II*\x00"))
    im = Image.open(p)
    pages = [np.array(pg, dtype=np.float32) for pg in ImageSequence.Iterator(im)]
    assert len(pages) == 3
    for i in range(3):
        assert pages[i].view(np.uint32).tobytes() == imgs[i].view(np.uint32).tobytes()
    tags = Image.open(p).tag_v2
    assert tags[305] == "RISA" and tags[258] == (32,) and tags[259] == 1 and tags[262] == 1
    assert tags[263] == 1 and tags[277] == 1 and tuple(tags[339]) == (3,)
    with pytest.raises(_lib.RisaError):
        _lib.check(lib.risa_tiff_write_f32(str(p).encode(), imgs.ctypes.data, 3, 40, 24, 9))


def test_backprojector_variant_shapes(risa):
    """The variant table the autotuner walks: every shape is well formed (a
    16-pixel-wide FP16 tile, whole pixel columns per thread, FP16 LDS planes
    within 32 KB), indices past the end are refused, and the kernel names
    the profiler reports follow from the shapes."""
    from risa_amd import _lib, device

    lib = _lib.load()
    for f16 in (0, 1):
        shapes = []
        while True:
            s = (C.c_int * 7)()
            rc = lib.risa_hip_bp_variant_shape(f16, len(shapes), s)
            if rc:
                assert rc == 1
                break
            shapes.append(tuple(s))
        assert len(shapes) >= 4
        assert len(set(shapes)) == len(shapes)
        for tx, ty, ppt, e, u, g, amax in shapes:
            assert ty % ppt == 0 and tx * ty // ppt <= 1024 and (tx * ty // ppt) % 64 == 0
            assert e >= 1 and u >= 1 and 1 <= amax <= 64
            if f16:
                assert tx == 16 and g in (1, 2, 4, 8)
                assert g * 2 * e * (tx * ty // ppt) * 16 <= 32 * 1024
            else:
                assert g in (1, 2, 4)   # slices per workgroup
        name = device.bp_kernel_name(bool(f16), 0)
        tx, ty, ppt, e, u, g, _ = shapes[0]
        assert name.endswith(f"<{tx}, {ty}, {ppt}, {e}, {u}, {g}>")
    assert lib.risa_hip_bp_variant_shape(0, -1, (C.c_int * 7)()) == 1


def test_synth_is_deterministic(risa, geom_c0):
    from risa_amd import synth

    a = synth.fan_frame(geom_c0, 0, 5, 9)
    b = synth.fan_frame(geom_c0, 0, 5, 9)
    c = synth.fan_frame(geom_c0, 0, 5, 10)
    assert a.dtype == np.uint16 and a.shape == geom_c0.fan_shape
    np.testing.assert_array_equal(a, b)
    assert (a != c).any()
    # reference frames: Poisson(60000) + N(200, 5) per sample
    ref = synth.reference_frames(geom_c0, 2, seed=3)
    m = ref.astype(np.float64).mean()
    assert abs(m - 60200) < 50
    assert abs(ref.astype(np.float64).std() - np.sqrt(60000 + 25)) < 10


def test_synth_phantom_statistics(risa):
    """Seeded ellipse phantoms: 10-20 ellipses, mu in [0, 0.05]/mm, inside the
    field of view (0.8 W/2 centres plus 0.4 W/2 semi-axes stay within 1.2 W/2)."""
    from risa_amd import synth

    for seed in range(5):
        e = synth.phantom(seed, 190.0)
        assert 10 <= len(e) <= 20
        cx, cy, a, b, mu = e[:, 0], e[:, 1], e[:, 2], e[:, 3], e[:, 5]   # {cx, cy, a, b, phi, mu}
        assert (mu >= 0).all() and (mu <= 0.05).all()
        assert (np.hypot(cx, cy) <= 0.8 * 95.0 + 1e-9).all()
        assert (a >= 0.05 * 95.0 - 1e-9).all() and (a <= 0.4 * 95.0 + 1e-9).all()
        assert (b >= 0.05 * 95.0 - 1e-9).all() and (b <= 0.4 * 95.0 + 1e-9).all()
    img = synth.phantom_image(1, 190.0, 64)
    assert img.shape == (64, 64) and img.max() > 0 and img.min() >= 0


def test_glados_framework_selftest(risa):
    """The C++ pipeline framework and the stages' host side, run on the CPU:
    queue bound and order, pool backpressure, sentinel propagation through a
    Source -> Stage -> Stage -> Sink pipeline, the module-major reference file
    layout (Attenuation.cu:275-321) and stage configuration errors."""
    import subprocess

    exe = ROOT / "risa_amd" / "bin" / "glados_selftest"
    assert exe.exists(), "build with `make -C risa_amd`"
    r = subprocess.run([str(exe), str(ROOT / "risa_amd" / "configs" / "c0.cfg")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "checks passed" in r.stdout
