"""GPU side of the file path: the device ingest kernel against the ingest
restatement (bit-exact), and the design-level writers against the reference's
writers fed with the same device fields."""
import os

import numpy as np
import pytest

from oracle import ingest
from poisson_caustic_design_amd import capi

pytestmark = pytest.mark.gpu
GOLDEN_IMG = os.path.join(os.path.dirname(__file__), "golden", "img")


@pytest.mark.parametrize("name,M", [("hello.png", 100), ("siggraph.png", 64), ("hello.png", 33)])
def test_gray_resize_bit_exact(ctx, name, M):
    rgba = capi.png_read_rgba8(os.path.join(GOLDEN_IMG, name))
    h, w = rgba.shape[:2]
    gw, gh, _, _, _ = ingest.driver_geometry(w, h, M, 1.0)
    got = capi.image_gray_resize(ctx, rgba, gw, gh)
    exp = ingest.resize(ingest.gray(rgba), gw, gh)
    assert got.shape == exp.shape and np.array_equal(got, exp)


def test_design_outputs_match_reference_writers(ctx, reflib, tmp_path):
    rgba = capi.png_read_rgba8(os.path.join(GOLDEN_IMG, "hello.png"))
    h, w = rgba.shape[:2]
    M = 16
    gw, gh, mx, my, mh = ingest.driver_geometry(w, h, M, 1.0)
    img = capi.image_gray_resize(ctx, rgba, gw, gh)
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(mx, my)
    cd.set_domain_resolution(gw, gh)
    cd.set_mesh_size(1.0, mh)
    cd.set_lens_focal_length(1.5)
    cd.set_lens_thickness(0.2)
    cd.initialize_solvers(img)
    cd.run(4, float(np.float32(0.01)), 3)
    src, tgt = cd.get("source_points"), cd.get("target_points")
    inv = cd.get("inverse_map")
    cases = [
        ("src.obj", lambda p: cd.save_solid_obj_source(p),
         lambda p: reflib.save_solid_obj(p, src, src, mx, my, 0.2, 1.0, mh)),
        ("tgt.obj", lambda p: cd.save_solid_obj_target(p),
         lambda p: reflib.save_solid_obj(p, tgt, src, mx, my, 0.2, 1.0, mh)),
        ("param.svg", lambda p: cd.export_paramererization_to_svg(p, 0.5),
         lambda p: reflib.export_svg(p, tgt, mx, my, 1.0, mh, 0.5)),
        ("inv.svg", lambda p: cd.export_inverted_transport_map(p, 1.0),
         lambda p: reflib.export_svg(p, inv, mx, my, 1.0, mh, 1.0)),
        ("tri.svg", lambda p: cd.export_to_svg(p, 1.0),
         lambda p: reflib.export_svg(p, tgt, mx, my, 1.0, mh, 1.0, triangles=True)),
    ]
    for name, ours, ref in cases:
        a, b = str(tmp_path / ("a_" + name)), str(tmp_path / ("b_" + name))
        ours(a)
        ref(b)
        assert open(a, "rb").read() == open(b, "rb").read(), name
    j = tmp_path / "h.json"
    cd.save_heightmap_json(str(j))
    assert j.read_text() == ingest.heightmap_json(cd.get("h"))
    cd.close()


def test_cli_matches_library_calls(ctx, tmp_path):
    """The command-line drop-in (main.cpp:137-272 flow) writes the same
    output.obj / heightmap.json / progress SVGs as the same calls made through
    the library directly."""
    import subprocess

    cli = os.path.join(os.path.dirname(capi.LIB_PATH), "caustic_design_amd")
    out, prog = tmp_path / "out", tmp_path / "prog"
    out.mkdir()
    prog.mkdir()
    png = os.path.join(GOLDEN_IMG, "siggraph.png")
    r = subprocess.run([cli, "--input_png", png, "--res_w", "24", "--output", str(out) + "/",
                        "--progress_out", str(prog) + "/", "--conv_tres=0.02"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    assert "Height solver done! Exporting as solidified obj" in r.stdout
    n_steps = r.stdout.count("starting iteration")
    # the same flow through the library
    rgba = capi.png_read_rgba8(png)
    h, w = rgba.shape[:2]
    gw, gh, mx, my, mh = ingest.driver_geometry(w, h, 24, 1.0)
    cd = capi.CausticDesign(ctx)
    cd.set_mesh_resolution(mx, my)
    cd.set_domain_resolution(gw, gh)
    cd.set_mesh_size(1.0, mh)
    cd.set_lens_focal_length(1.5)
    cd.set_lens_thickness(0.2)
    cd.set_first_hit(1)  # as the CLI: the reference BVH's first hit
    cd.initialize_solvers(capi.image_gray_resize(ctx, rgba, gw, gh))
    conv = float(np.float32(0.02))
    steps = 0
    for _ in range(50):
        steps += 1
        if cd.perform_transport_iteration() < conv:
            break
    for i in range(3):
        cd.perform_height_map_iteration(i)
    assert steps == n_steps
    cd.save_solid_obj_source(str(tmp_path / "api.obj"))
    cd.save_heightmap_json(str(tmp_path / "api.json"))
    cd.export_paramererization_to_svg(str(tmp_path / "api.svg"), 1.0)
    assert (out / "output.obj").read_bytes() == (tmp_path / "api.obj").read_bytes()
    assert (out / "heightmap.json").read_bytes() == (tmp_path / "api.json").read_bytes()
    assert (prog / f"parameterization_{n_steps}.svg").read_bytes() == \
        (tmp_path / "api.svg").read_bytes()
    assert (prog / "parameterization_0.svg").exists() and (prog / "inverted.svg").exists()
    cd.close()
