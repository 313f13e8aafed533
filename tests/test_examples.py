"""Every case file of the reference's example/flow runs through the case
driver's handlers (host only, tclbx_case_check), or names a NodeType that
only models outside the engine declare.  For a file that runs there is a
model of the engine for which every element becomes its handler (the
reference's HandlerFactory set, src/HandlerFactory.h:8-17) with no error,
its units, settings, zones and geometry resolve, and its Solve elements step
through their iterations and callbacks.  A file of the other kind fails, for
every model tried, at its geometry with the reference's message for a
NodeType the model lacks ("Unknown flag (in xml): <name>",
src/Geometry.cpp.Rt), and no model of the engine of the file's dimension
declares that NodeType (e.g. Inlet/Outlet in 3D: d3q19_adj and
d3q19_heat_adj_art).  The files are
read where they lie in the reference's tree, so the test runs where that
tree is present."""
import glob
import os
import re

import pytest

import tclb_amd

REF_ROOT = "/root/reference"
EXAMPLES = os.path.join(REF_ROOT, "example", "flow")
MODELS = ["d2q9_SRT", "d2q9", "d3q19", "d3q27", "d3q27_cumulant", "d3q27_cumulant_AVG", "d3q27_cumulant_IB"]


def _examples():
    return sorted(glob.glob(os.path.join(EXAMPLES, "*", "*.xml")))


def _declared_anywhere(nodetype, two_d):
    """Whether any model of the engine with the file's dimension declares
    the NodeType."""
    for m in tclb_amd.models():
        if m.startswith("d2q9") != two_d:
            continue
        try:
            tclb_amd.Model(m).nodetype(nodetype)
            return True
        except tclb_amd.TclbxError:
            pass
    return False


@pytest.mark.skipif(not _examples(), reason="the reference's example case files are not present")
def test_every_example_runs_through_the_handlers(monkeypatch, tmp_path):
    monkeypatch.chdir(REF_ROOT)   # STL paths in the files are relative to the reference's root
    files = _examples()
    assert len(files) >= 17
    ran, foreign = {}, {}
    for f in files:
        name = os.path.relpath(f, EXAMPLES)
        errors = {}
        for model in MODELS:
            try:
                c = tclb_amd.Case(model, f, output_root=str(tmp_path) + "/")
            except tclb_amd.TclbxError as e:
                errors[model] = str(e)
                continue
            if model.startswith("d2q9") != (c.nz == 1):
                continue
            try:
                c.check()
            except tclb_amd.TclbxError as e:
                errors[model] = str(e)
                continue
            bad = [m for m in c.messages if m.startswith("ERROR")]
            if bad:
                errors[model] = "; ".join(bad)
                continue
            ran.setdefault(name, []).append(model)
        if name in ran:
            continue
        lacking = {re.search(r"Unknown flag \(in xml\): (\w+)", e).group(1) if "Unknown flag" in e else None
                   for e in errors.values() if "is a 2D model" not in e and "2D model" not in e}
        assert None not in lacking and lacking, (name, errors)
        for t in lacking:
            assert not _declared_anywhere(t, name.startswith("2d")), (name, t)
        foreign[name] = sorted(lacking)
    print("ran:", ran)
    print("NodeTypes of other models:", foreign)
    assert len(ran) >= 10
    assert len(ran) + len(foreign) == len(files)


def test_check_reports_an_unknown_element(tmp_path):
    """An element no handler takes is reported as the reference reports it
    and passed over (NullHandler); a Param naming no setting fails."""
    xml = """<CLBConfig output="out/"><Geometry nx="16" ny="8"><BGK><Box/></BGK></Geometry>
        <Model><Param name="nu" value="0.1"/></Model><Frobnicate/><Solve Iterations="10"/></CLBConfig>"""
    c = tclb_amd.Case("d2q9_SRT", xml=xml, output_root=str(tmp_path) + "/").check()
    assert "ERROR: Unknown Handler: Frobnicate" in c.messages
    bad = """<CLBConfig><Geometry nx="16" ny="8"><BGK><Box/></BGK></Geometry>
        <Model><Param name="NoSuchSetting" value="1"/></Model><Solve Iterations="10"/></CLBConfig>"""
    with pytest.raises(tclb_amd.TclbxError):
        tclb_amd.Case("d2q9_SRT", xml=bad, output_root=str(tmp_path) + "/").check()
