"""The C++ `Simulation` subclass that INTEGRATION.md shows a maintainer adding
compiles against include/garm_gpu.h.

The sketch is taken from INTEGRATION.md as written (its first C++ block) and
compiled, with every member function instantiated, against a stand-in for
the reference's `Simulation` interface (Common/Util/Simulation.h:10-33: the
same virtual functions and signatures, so every `override` in the sketch
must match one) and empty stand-ins for the reference's Yaml.h and
ImplicitSurface.h.  The option snippets that follow it are compiled as the
body of a member function of the same class, so a misspelt option or a
wrong call signature anywhere in the document fails here.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SIMULATION_H = r"""
#pragma once
#include "Yaml.h"
#include <filesystem>
#include <stdexcept>
#include <string>
#include <sys/types.h>
#include <vector>
namespace PhysX {
class Simulation
{
protected:
	double _time = 0;
public:
	Simulation() = default;
	Simulation(const Simulation &rhs) = delete;
	Simulation &operator=(const Simulation &rhs) = delete;
	virtual ~Simulation() = default;
	void setTime(const double time) { _time = time; }
	virtual double getTimeStep(const uint frameRate, const double stepRate) const { return 1.0 / frameRate / stepRate; }
	virtual int dimension() const = 0;
	virtual void writeDescription(YAML::Node &root) const = 0;
	virtual void writeFrame(const std::string &frameDir, const bool staticDraw) const = 0;
	virtual void saveFrame(const std::string &frameDir) const = 0;
	virtual void loadFrame(const std::string &frameDir) = 0;
	virtual void initialize() { }
	virtual void advance(const double dt) = 0;
};
}
"""


def _cpp_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, flags=re.S)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_integration_sketch_compiles(tmp_path):
    blocks = _cpp_blocks()
    assert len(blocks) >= 3, "INTEGRATION.md lost its C++ blocks"
    cls, options = blocks[0], blocks[1:]
    (tmp_path / "Simulation.h").write_text(SIMULATION_H)
    (tmp_path / "Yaml.h").write_text("#pragma once\nnamespace YAML { struct Node {}; }\n")
    (tmp_path / "ImplicitSurface.h").write_text("#pragma once\n")
    # the option snippets as a member function of the sketched class (they
    # refer to its members _opts, _ctx and a surface-tension options member _st)
    body = "\n".join(options)
    extra = ("\tgarm_surface_tension_opts _st{};\n"
             "\tunsigned long long seed = 1;\n"
             "\tvoid configure() {\n" + body + "\n\t}\n")
    assert cls.rstrip().endswith("}"), "unexpected end of the class block"
    # insert the member function before the class's closing "};"
    pos = cls.rfind("};")
    src = cls[:pos] + extra + cls[pos:]
    src += "\ntemplate class PhysX::GpuLevelSetLiquid<2>;\ntemplate class PhysX::GpuLevelSetLiquid<3>;\n"
    (tmp_path / "sketch.cpp").write_text(src)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wno-unused-variable", "-I", str(tmp_path),
                        "-I", os.path.join(ROOT, "include"), str(tmp_path / "sketch.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
