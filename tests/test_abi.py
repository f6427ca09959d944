"""The C ABI: exported symbols, struct layouts and host-side error behaviour."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import nanogi_amd as ngi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ngi.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(ngi_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    decl = _declared_functions()
    assert len(decl) >= 25
    lib = ngi.lib()
    missing = [n for n in decl if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(decl) == sorted(ngi.EXPORTS)


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof from a C compiler against the ctypes mirror."""
    structs = {"ngi_primitive": ngi.Primitive, "ngi_scene_desc": ngi.SceneDesc, "ngi_bvh_opts": ngi.BvhOpts,
               "ngi_bvh_info": ngi.BvhInfo, "ngi_ray": ngi.Ray, "ngi_hit": ngi.Hit,
               "ngi_render_params": ngi.RenderParams, "ngi_stats": ngi.Stats, "ngi_batch_opts": ngi.BatchOpts,
               "ngi_tune_trial": ngi.TuneTrial}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    for cname, cls in structs.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert int(out[f"{cname}.{field}"]) == getattr(cls, field).offset, f"{cname}.{field}"


def test_renderer_strings():
    # RendererType_String (src/nanogi.cpp:61-69)
    for i, name in enumerate(["pt", "ptdirect", "lt", "ltdirect", "bdpt", "ptmnee"]):
        assert ngi.renderer_from_string(name) == i
    assert ngi.renderer_from_string("PT") == -1
    assert ngi.renderer_from_string("path") == -1


def test_default_params_are_reference_defaults():
    p = ngi.default_params()
    # src/nanogi.cpp:2004-2007, RR after every intersection with p = 0.5 (583-591)
    assert (p.width, p.height, p.num_samples, p.max_num_vertices) == (1280, 720, 10000000, -1)
    assert p.renderer == ngi.RENDERER_PT and p.sampling == ngi.SAMPLING_REFERENCE_UNIFORM
    assert (p.rr_depth, p.rr_prob) == (1, 0.5)
    assert p.engine == ngi.ENGINE_PERSISTENT and p.precision == ngi.PRECISION_FP32_FP64ACC


def test_unsupported_renderers_fail_before_touching_the_device():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    for r in ("ptmnee",):
        with pytest.raises(ngi.NgiError) as e:
            ngi.render(s, renderer=r, width=8, height=8)
        assert e.value.status == ngi.ERR_UNSUPPORTED
    with pytest.raises(ngi.NgiError):
        ngi.default_params(renderer="nosuch")


def test_invalid_render_arguments():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    bad = [dict(width=0), dict(height=-1), dict(sampling=7), dict(precision=9), dict(engine=5),
           dict(rr_depth=1, rr_prob=0.0), dict(max_num_vertices=-2),
           dict(sampling=ngi.SAMPLING_REFERENCE_UNIFORM, num_samples=0)]
    for kw in bad:
        with pytest.raises(ngi.NgiError) as e:
            ngi.render(s, **dict(dict(width=8, height=8), **kw))
        assert e.value.status == ngi.ERR_INVALID_ARG, kw


def test_scene_without_sensor_is_rejected():
    from tests import scenes
    B = scenes.SceneBuilder()
    p = B.prim(type=ngi.PRIM_D, D_R=(1, 1, 1))
    B.rect(p, (0, 0, 0), (0, 1, 0), (1, 0, 0), 1, 1)
    s = B.scene()
    with pytest.raises(ngi.NgiError) as e:
        ngi.render(s, width=8, height=8)
    assert e.value.status == ngi.ERR_INVALID_SCENE


def test_adjoint_rejects_light_tracers_and_bdpt():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    for r in ("lt", "ltdirect", "bdpt"):
        p = ngi.default_params(renderer=r, width=4, height=4, sampling=ngi.SAMPLING_PER_PIXEL, spp=1)
        with pytest.raises(ngi.NgiError) as e:
            ngi.render_adjoint(s, p, np.zeros((4, 4, 3)))
        assert e.value.status == ngi.ERR_UNSUPPORTED


def test_bdpt_rejects_fp64_shading():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    with pytest.raises(ngi.NgiError) as e:
        ngi.render(s, renderer="bdpt", width=4, height=4, precision=ngi.PRECISION_FP64_ALL)
    assert e.value.status == ngi.ERR_UNSUPPORTED


def test_autotune_arguments_are_checked_before_the_device():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    ok = dict(width=8, height=8, spp=1, sampling=ngi.SAMPLING_PER_PIXEL)
    with pytest.raises(ngi.NgiError) as e:
        ngi.autotune(s, ngi.default_params(**ok), tune_paths=-1)
    assert e.value.status == ngi.ERR_INVALID_ARG
    for kw, status in [(dict(width=0), ngi.ERR_INVALID_ARG), (dict(renderer="ptmnee"), ngi.ERR_UNSUPPORTED),
                       (dict(renderer="bdpt", precision=ngi.PRECISION_FP64_ALL), ngi.ERR_UNSUPPORTED)]:
        p = ngi.default_params(**dict(ok, **kw))
        with pytest.raises(ngi.NgiError) as e:
            ngi.autotune(s, p)
        assert e.value.status == status, kw
        assert "autotune" in e.value.message


def test_no_gpu_reports_device_error():
    if ngi.device_count() > 0:
        pytest.skip("a GPU is present")
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    with pytest.raises(ngi.NgiError) as e:
        s.trace(ngi.make_rays([[0.5, 0.5, 2]], [[0, 0, -1]]))
    assert e.value.status == ngi.ERR_DEVICE
    with pytest.raises(ngi.NgiError) as e:
        ngi.render(s, width=8, height=8, spp=1, sampling=ngi.SAMPLING_PER_PIXEL)
    assert e.value.status == ngi.ERR_DEVICE
    assert "no HIP device" in e.value.message


def test_batch_arguments_are_checked_before_the_device():
    a = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    b = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 2, 1.0, 200)
    ok = dict(width=8, height=8, spp=1, sampling=ngi.SAMPLING_PER_PIXEL)
    cases = [
        ([dict(ok, renderer="pt"), dict(ok, renderer="ptdirect")], ngi.ERR_UNSUPPORTED),
        ([dict(ok), dict(ok, precision=ngi.PRECISION_FP64_ALL)], ngi.ERR_UNSUPPORTED),
        ([dict(ok), dict(ok, renderer="bdpt")], ngi.ERR_UNSUPPORTED),
        ([dict(ok), dict(ok, width=0)], ngi.ERR_INVALID_ARG),
        ([dict(ok), dict(ok, spp=0)], ngi.ERR_INVALID_ARG),
        ([dict(ok), dict(ok, rr_depth=1, rr_prob=2.0)], ngi.ERR_INVALID_ARG),
    ]
    for kws, status in cases:
        with pytest.raises(ngi.NgiError) as e:
            ngi.render_batch([a, b], [ngi.default_params(**k) for k in kws])
        assert e.value.status == status, kws
    with pytest.raises(ngi.NgiError) as e:
        ngi.render_batch([], [])
    assert e.value.status == ngi.ERR_INVALID_ARG
    with pytest.raises(ValueError):
        ngi.render_batch([a], [])
    # The reusable batch validates the same way, and its options too.
    with pytest.raises(ngi.NgiError) as e:
        ngi.Batch([a, b], [ngi.default_params(**dict(ok, renderer="pt")), ngi.default_params(**dict(ok, renderer="ptdirect"))])
    assert e.value.status == ngi.ERR_UNSUPPORTED
    o = ngi.batch_opts()
    assert (o.block_size, o.blocks_per_cu, o.use_graph, o.chunk_paths, o.tune_paths) == (0, 0, 1, 0, 1 << 24)
    with pytest.raises(ngi.NgiError) as e:
        ngi.Batch([a], [ngi.default_params(**ok)], ngi.batch_opts(block_size=100))
    assert e.value.status == ngi.ERR_INVALID_ARG
    with pytest.raises(ngi.NgiError) as e:
        ngi.Batch([a], [ngi.default_params(**ok)], ngi.batch_opts(chunk_paths=-1))
    assert e.value.status == ngi.ERR_INVALID_ARG
