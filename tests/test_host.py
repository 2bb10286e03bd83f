"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/*.h declares, the ctypes structures match the C
layout, and the host mirror of the reference types behaves like the
reference's own tests say."""
import ctypes
import glob
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from ethersph import _lib
from ethersph import particles as P
from ethersph import scenarios as S
from ethersph.time_discretization import (DensityReinitializedFixedTimeIntervalForwardEuler, isCheckStep,
                                          isDensityReinitializedStep, isOutputStep)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tests.json")))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(esph_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    """Every declared entry point has a ctypes signature in the binding."""
    src = open(os.path.join(ROOT, "ethersphcpu.jl_amd", "ethersph", "_lib.py")).read()
    missing = [n for n in declared_symbols() if f'"{n}"' not in src]
    assert not missing, missing


def test_struct_layout_matches_c():
    code = r'''
#include <stddef.h>
#include <stdio.h>
#include "ethersph.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(esph_scene_desc), sizeof(esph_fluid_arrays), sizeof(esph_wall_arrays));
  printf("%zu %zu %zu %zu %zu\n", offsetof(esph_scene_desc, mu_0), offsetof(esph_scene_desc, dt),
         offsetof(esph_scene_desc, check_max), offsetof(esph_scene_desc, domain_max),
         offsetof(esph_scene_desc, fluid_capacity));
  printf("%zu %zu\n", offsetof(esph_fluid_arrays, cp), offsetof(esph_wall_arrays, kappa));
  printf("%zu %zu %zu %zu\n", offsetof(esph_scene_desc, io_axis), offsetof(esph_scene_desc, profile_center),
         offsetof(esph_scene_desc, verlet_skin), offsetof(esph_scene_desc, check_planes));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(code)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    vals = list(map(int, out))
    SD, FA, WA = _lib.SceneDesc, _lib.FluidArrays, _lib.WallArrays
    assert vals[:3] == [ctypes.sizeof(SD), ctypes.sizeof(FA), ctypes.sizeof(WA)]
    assert vals[3:8] == [SD.mu_0.offset, SD.dt.offset, SD.check_max.offset, SD.domain_max.offset,
                         SD.fluid_capacity.offset]
    assert vals[8:10] == [FA.cp.offset, WA.kappa.offset]
    assert vals[10:] == [SD.io_axis.offset, SD.profile_center.offset, SD.verlet_skin.offset,
                         SD.check_planes.offset]


def test_particle_constructors_zero_initialise():
    g = GOLD["particle_zero_init"]
    assert P.CompulsiveWallParticle(2).gap_ == g["CompulsiveWallParticle.gap_"]
    assert P.StaticVelocityWallParticle(2).v_vec_.tolist() == g["StaticVelocityWallParticle.v_vec_"]
    assert P.ThermostaticWallParticle(2).t_ == g["ThermostaticWallParticle.t_"]
    assert P.CommonLiquidParticle(2).mass_ == g["CommonLiquidParticle.mass_"]
    assert P.ThermalLiquidParticle(2).t_ == g["ThermalLiquidParticle.t_"]


def test_particle_array_roundtrip():
    ps = [P.ThermalLiquidParticle(2) for _ in range(5)]
    for i, p in enumerate(ps):
        p.x_vec_[:] = [i, 2 * i]
        p.rho_ = 1000.0 + i
        p.t_ = 0.1 * i
    a = P.LiquidParticles.from_particles(ps)
    assert a.thermal and a.x_vec_.shape == (5, 2) and a.rho_[3] == 1003.0 and a.t_[4] == 0.4
    a.rho_ += 1
    a.write_back(ps)
    assert ps[2].rho_ == 1003.0


def test_time_discretization_predicates_and_positional_order():
    td = DensityReinitializedFixedTimeIntervalForwardEuler(1e-3, 100, 10, 7, 30)
    assert td.check_interval_ == 7 and td.density_reinitialized_interval_ == 30
    assert isDensityReinitializedStep(60, td) and not isDensityReinitializedStep(61, td)
    assert isOutputStep(20, td) and isCheckStep(14, td)


def test_scenario_sizes():
    sc = S.collapse_dry_2d()
    assert (sc.n_fluid, sc.n_wall) == (5000, 1518)
    assert sc.kernel.h_ == pytest.approx(0.03) and sc.kernel.influence_radius_ == 3.0 * 0.02
    assert sc.model.c_0_ == pytest.approx(62.64184, rel=1e-6)
    sc = S.thermal_cavity_2d()
    assert (sc.n_fluid, sc.n_wall) == (40000, 2436) and sc.solver == _lib.SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL
    sc = S.lid_driven_cavity_2d()
    assert sc.solver == _lib.SOLVER_FLUID_VELOCITY


def test_scenario_3d_size():
    sc = S.dam_break_3d()
    assert sc.n_fluid == 1_013_040 and sc.n_wall == 645_804


def test_lattice_matches_reference_operation_order():
    """x = (x0 + (j - 1) dx) + dx / 2 exactly (createRectangleParticles)."""
    X = S.rectangle(0.02, 0.0, -0.06, 7, 3)
    for idx in range(21):
        i, j = divmod(idx, 7)
        assert X[idx, 0] == (0.0 + j * 0.02) + 0.02 / 2
        assert X[idx, 1] == (-0.06 + i * 0.02) + 0.02 / 2


def test_product_has_no_oracle_dependency():
    """The product package and the HIP sources never reference the oracle."""
    pkg = os.path.join(ROOT, "ethersphcpu.jl_amd")
    for f in glob.glob(os.path.join(pkg, "**", "*.*"), recursive=True):
        if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp", ".jl")):
            txt = open(f).read()
            assert "oracle" not in txt.lower(), f


def _julia_struct(src, name):
    """Field (name, Julia type) list of `struct name ... end` in the shim."""
    import re

    m = re.search(r"struct %s\n(.*?)\nend" % name, src, re.S)
    assert m, name
    out = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if line:
            f, t = line.split("::")
            out.append((f.strip(), t.strip()))
    return out


def test_julia_shim_matches_the_c_abi():
    """The Julia ccall binding (not runnable here: no Julia in the image) lays
    its structs out like include/ethersph.h -- same fields, order and widths
    as the ctypes binding, whose layout test_struct_layout_matches_c checks
    against the C compiler -- and calls only entry points the header declares."""
    import re

    src = open(os.path.join(ROOT, "ethersphcpu.jl_amd", "julia", "EtherSPHMI355X.jl")).read()
    jl_size = {"Int32": 4, "Int64": 8, "Float64": 8, "Ptr{Float64}": 8, "NTuple{3,Float64}": 24,
               "NTuple{32,Float64}": 256}
    for jname, cls in (("EsphSceneDesc", _lib.SceneDesc), ("EsphFluidArrays", _lib.FluidArrays),
                       ("EsphWallArrays", _lib.WallArrays)):
        jf = _julia_struct(src, jname)
        cf = cls._fields_
        assert [f for f, _ in jf] == [f for f, _ in cf], jname
        for (f, t), (_, ct) in zip(jf, cf):
            assert jl_size[t] == ctypes.sizeof(ct), (jname, f, t)
    called = set(re.findall(r"ccall\(\(:(esph_\w+), libesph\)", src))
    assert called and called <= set(declared_symbols()), called - set(declared_symbols())
