"""The C-ABI library loads, exports every symbol include/shkz_mi355x.h declares, and its struct
layouts agree with a C compiler's (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import shiokaze_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shkz_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shkz_[a-z_0-9]+)\s*\(", text)))


def test_library_loads():
    assert S.lib().shkz_abi_version() == S.ABI_VERSION


def test_every_declared_symbol_is_exported():
    L = S.lib()
    decl = declared_functions()
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(S.EXPORTS) == decl


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "shkz_mi355x.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",sizeof(shkz_grid),'
                   'sizeof(shkz_params),sizeof(shkz_result),sizeof(shkz_scene_desc),sizeof(shkz_kernel_stat),'
                   'offsetof(shkz_grid,scene_id),offsetof(shkz_result,t_projection),sizeof(shkz_server_config),'
                   'sizeof(shkz_server_stats),offsetof(shkz_result,t_latency),'
                   'offsetof(shkz_server_config,hbm_budget_bytes),sizeof(shkz_grid2d),'
                   'offsetof(shkz_grid2d,scene_id),offsetof(shkz_grid,active_out_w),sizeof(shkz_velocity_field),'
                   'offsetof(shkz_velocity_field,active_w),sizeof(shkz_advect_params),'
                   'offsetof(shkz_advect_params,background));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [ctypes.sizeof(S.Grid), ctypes.sizeof(S.Params), ctypes.sizeof(S.Result), ctypes.sizeof(S.SceneDesc),
            ctypes.sizeof(S.KernelStat), S.Grid.scene_id.offset, S.Result.t_projection.offset,
            ctypes.sizeof(S.ServerConfig), ctypes.sizeof(S.ServerStats), S.Result.t_latency.offset,
            S.ServerConfig.hbm_budget_bytes.offset, ctypes.sizeof(S.Grid2D), S.Grid2D.scene_id.offset,
            S.Grid.active_out_w.offset, ctypes.sizeof(S.VelocityField), S.VelocityField.active_w.offset,
            ctypes.sizeof(S.AdvectParams), S.AdvectParams.background.offset]
    assert got == want


def test_default_params_are_the_reference_defaults():
    p = S.default_params()
    assert p.residual == 1e-4 and p.max_iterations == 30000           # pcg.cpp:76-77
    assert p.eps_fluid == 1e-2 and p.eps_solid == 1e-2                # macutility3.cpp:409-410
    assert p.second_order_fluid == 1 and p.second_order_solid == 1    # macpressuresolver3.cpp:302-303
    assert p.gain == 1.0 and p.warm_start == 0 and p.norm == S.NORM_INF


def test_last_error_starts_empty():
    assert S.last_error() == ""


@pytest.mark.parametrize("kind", [S.SCENE_DROP, S.SCENE_DROP_SOLIDS, S.SCENE_SPHERES_BOX, S.SCENE_RANDOM_DROP])
def test_scene_synthesis_is_deterministic(kind):
    a = S.make_scene(kind, 20, seed=3)
    b = S.make_scene(kind, 20, seed=3)
    c = S.make_scene(kind, 20, seed=4)
    assert np.array_equal(a.fluid, b.fluid) and np.array_equal(a.u, b.u) and np.array_equal(a.wv, b.wv)
    assert not np.array_equal(a.u, c.u)
    frac = (a.fluid < 0).mean()
    assert 0.05 < frac < 0.6
    assert (a.solid is not None) == (kind in (S.SCENE_DROP_SOLIDS, S.SCENE_SPHERES_BOX))


def test_scene_velocity_statistics():
    sc = S.make_scene(S.SCENE_DROP, 32, seed=0, gravity=0.0)
    u = np.concatenate([sc.u.ravel(), sc.v.ravel(), sc.wv.ravel()])
    assert abs(u.mean()) < 0.01 and abs(u.std() - 1.0) < 0.01
    g = S.make_scene(S.SCENE_DROP, 32, seed=0)
    assert np.allclose(g.v - sc.v, np.float32(-9.8 / 120.0))


def test_scene_liquid_is_pushed_out_of_solids():
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 32, seed=1)
    # cell centres inside a solid are never liquid
    s_cell = sc.solid[:-1, :-1, :-1]  # nodal sample at the cell's low corner, inside test is loose
    assert not ((sc.fluid < 0) & (s_cell < -2.0 / 32)).any()


def test_smoke_scene_is_all_fluid():
    """SCENE_SMOKE: every cell liquid (macsmoke3.cpp:106), a consistent pure-Neumann system."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    from oracle import oracle as O
    sc = S.make_scene(S.SCENE_SMOKE, 24, seed=3)
    assert (sc.fluid == -1.0).all() and sc.solid is not None and (sc.solid < 0).any()
    A = O.assemble(sc)
    M = sp.csr_matrix((A["val"], A["col"], A["rowptr"]), shape=(A["n"], A["n"]))
    nc, lab = connected_components(M, directed=False)
    for c in range(nc):  # consistent: the rhs sums to zero on every component
        assert abs(A["b"][lab == c].sum()) <= 1e-9 * np.abs(A["b"]).sum()
    assert np.allclose(np.asarray(M.sum(axis=1)).ravel(), 0.0, atol=1e-9)  # no Dirichlet term
    assert (S.make_scene2d(S.SCENE_SMOKE, 30, 20, seed=3).fluid == -1.0).all()


@pytest.mark.skipif(not os.path.isdir("/root/reference/include"), reason="Shiokaze tree not present")
@pytest.mark.parametrize("module", ["macpressuresolver3_mi355x", "macpressuresolver2_mi355x", "pcg_mi355x"])
def test_shiokaze_adapter_compiles_against_reference_headers(module):
    """The drop-in module sources type-check against the reference's own interface headers."""
    src = os.path.join(ROOT, "shiokaze_amd", "host", module + ".cpp")
    subprocess.check_call(["g++", "-std=c++14", "-fsyntax-only", "-DNUM_THREAD=8", "-DINDEX_TYPE=size_t",
                           "-DFLOAT_TYPE=double", "-I/root/reference/include",
                           "-I/root/reference/local/include", src])


def test_host_harness_is_built():
    assert os.path.exists(os.path.join(ROOT, "shiokaze_amd", "lib", "shkz_harness"))
