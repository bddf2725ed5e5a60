"""The Lua-subset reader of the problem files (pfem_amd.lua_config): Lua's semantics on the
constructs the reference's examples use, the SolTable interface (SolTable.hpp:10-117), and,
when the reference's example directory is present in this container, that every example
problem file loads and its initStates / BC functions return vectors."""
import glob
import math
import os

import numpy as np
import pytest

from pfem_amd.lua_config import LuaError, LuaState, SolTable

SRC = r'''
-- a problem in the examples' style (written for this test)
Problem = {
    id = "IncompNewtonNoT",
    simulationTime = 4,
    verboseOutput = false,
    Mesh = { hchar = 0.02, alpha = 1.2, boundingBox = {-0.01, -1, 1.5, 100}, exclusionZones = {},
             mshFile = "geometry.msh" },
    Material = { mu = 1e-3, rho = 1000, gamma = 0 },
    IC = { },
    Solver = {
        id = "PSPG", maxDT = 1e-3, adaptDT = true,
        MomContEq = { minRes = 1e-6, maxIter = 10, bodyForce = {0, -9.81}, residual = "Ax_f", BC = { } }
    }
}

function Problem.IC:initStates(pos)
    local g = -Problem.Solver.MomContEq.bodyForce[2]
    local rho = Problem.Material.rho
    local p = 0
    if(pos[2] <= 0.5 and pos[1] <= 0.3 + 2*Problem.Mesh.hchar) then
        p = rho*g*(0.5 - pos[2])
    elseif pos[2] > 10 then
        p = -1
    else
        p = 0
    end
    return {0, 0, p}
end

function Problem.Solver.MomContEq.BC:WallV(pos, t)
    local DT = 0.5
    if t < DT then
        return {math.sin(math.pi*t/DT)*(1 - pos[2]^2), 0}
    end
    return {0, 0}
end

Problem.IC.BoundaryFixed = true
local helper = 2^3^2
Problem.Mesh.checks = {helper, -2^2, 7 % 3, -7 % 3, 1 .. "x", (nil or 4), (false and 1), (3 and 5),
                       math.max(1, 9, 4), math.abs(-2.5), #{1, 2, 3}, 10/4}
'''


def test_tables_functions_and_solTable():
    st = LuaState.from_string(SRC)
    prob = SolTable("Problem", st)
    assert prob.getName() == "Problem"
    assert prob.checkAndGet("id") == "IncompNewtonNoT"
    assert prob.checkAndGet("verboseOutput") is False
    mesh = SolTable("Mesh", prob)
    assert mesh.checkAndGet("hchar") == 0.02
    assert mesh.checkAndGet("boundingBox") == [-0.01, -1, 1.5, 100]
    assert mesh.checkAndGet("exclusionZones") == []
    assert mesh.doesVarExist("alpha") and not mesh.doesVarExist("omega")
    with pytest.raises(LuaError, match="omega was not found in table Mesh"):
        mesh.checkAndGet("omega")
    with pytest.raises(LuaError, match="Could not find table Nope"):
        SolTable("Nope", prob)
    ic = SolTable("IC", prob)
    assert ic.checkAndGet("BoundaryFixed") is True
    # initStates(pos): a method (self = the IC table), globals reached from inside
    assert ic.call("initStates", [0.1, 0.2]) == [0, 0, pytest.approx(1000 * 9.81 * 0.3)]
    assert ic.call("initStates", np.array([0.5, 0.2])) == [0, 0, 0]
    assert ic.call("initStates", [0.1, 11.0]) == [0, 0, -1]
    bc = SolTable("BC", SolTable("MomContEq", SolTable("Solver", prob)))
    v = bc.call("WallV", [0.0, 0.5], 0.25)
    assert v[0] == pytest.approx(math.sin(math.pi * 0.5) * 0.75) and v[1] == 0
    assert bc.call("WallV", [0.0, 0.5], 0.75) == [0, 0]
    assert bc.checkCallNoThrow("WallV", [0, 0], 0.0) and not bc.checkCallNoThrow("FreeV", [0, 0], 0.0)
    with pytest.raises(LuaError, match="Cannot find method FreeV"):
        bc.checkCall("FreeV", [0, 0], 0.0)
    # Lua semantics: ^ right-associative and above unary minus, % floored, .., and/or values
    assert mesh.checkAndGet("checks") == [512, -4, 1, 2, "1x", 4, False, 5, 9, 2.5, 3, 2.5]
    seen = []
    prob.for_each(lambda k, v: seen.append(k))
    assert {"id", "Mesh", "Material", "Solver", "IC"} <= set(seen)


def test_errors_are_reported():
    with pytest.raises(LuaError):
        LuaState.from_string("x = = 1")
    with pytest.raises(LuaError, match="arithmetic"):
        LuaState.from_string("x = {} + 1")
    with pytest.raises(LuaError, match="outside the supported subset"):
        LuaState.from_string("while true do end")


EXAMPLES = "/root/reference/examples"


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="the reference's examples are not in this container")
def test_reference_example_problem_files_load():
    files = sorted(glob.glob(os.path.join(EXAMPLES, "*", "*", "*.lua")))
    assert len(files) >= 40
    problems = 0
    for path in files:
        st = LuaState.from_file(path)
        if st["Problem"] is None:  # helper scripts of other problems (functions only)
            continue
        problems += 1
        prob = SolTable("Problem", st)
        dim = 3 if "/3D/" in path else 2
        assert prob.checkAndGet("id")
        mesh = SolTable("Mesh", prob)
        assert len(mesh.checkAndGet("boundingBox")) == 2 * dim
        ic = SolTable("IC", prob)
        if ic.doesVarExist("initStates"):
            states = ic.call("initStates", [0.1] * dim)
            assert isinstance(states, list) and len(states) >= 1 and all(
                isinstance(x, (int, float)) for x in states), (path, states)
    assert problems >= 40
