"""Extract the golden vectors of elPaSo-Core's own unit tests into JSON fixtures.

Run once in a container that holds the elPaSo-Core sources at
/root/reference (GPU runs read only the committed JSON):

    python tests/golden/make_golden.py

It reads numbers only (the reference arrays of the gtest headers and the
2-element Fluid8 HDF5 fixture through h5dump's text output) and writes
tests/golden/elpaso_golden.json.  Nothing of the reference is executed.
"""
import json
import os
import re
import subprocess

REF = "/root/reference/elpasoCore/test"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elpaso_golden.json")


def _floats(txt):
    return [float(v) for v in re.findall(r"-?\d+(?:\.\d+)?(?:e[-+]?\d+)?", txt)]


def _array(src, name, after=0):
    i = src.index(name, after)
    j = src.index("{", i)
    k = src.index("}", j)
    return _floats(src[j + 1:k]), k


def main():
    g = {}
    hex8 = open(os.path.join(REF, "shape/testshapehex8.h")).read()
    vol, k = _array(hex8, "refData[256]")
    face, _ = _array(hex8, "refData[48]", k)
    g["hex8_table"] = {"source": "test/shape/testshapehex8.h:71-105", "tol": 1e-6, "values": vol}
    g["hex8_face_table"] = {"source": "test/shape/testshapehex8.h:194-203", "tol": 1e-6, "values": face}
    faces = []
    for f in range(6):
        m = re.search(r"nodesFace%d\{([^}]*)\}" % f, hex8)
        faces.append([int(v) for v in m.group(1).split(",")])
    g["hex8_face_nodes"] = {"source": "test/shape/testshapehex8.h:141-154", "values": faces}

    sf = open(os.path.join(REF, "shape/testshapefunction.h")).read()
    for func, key, names in (("testShapeFunctionLagrangeHex20", "hex20_at_gp4", ("refValFun", "refValXi", "refValEta", "refValZeta")),
                             ("testShapeFunctionLagrangeHex8()", "hex8_at_gp4", ("refValFun", "refValXi", "refValEta", "refValZeta")),
                             ("testShapeFunctionSerendipityQuad8", "quad8_at_gp4", ("refValFun", "refValXi", "refValEta"))):
        i = sf.index(func)
        d = {}
        for n in names:
            vals, _ = _array(sf, n, i)
            d[n] = vals
        g[key] = {"source": f"test/shape/testshapefunction.h ({func})", "tol": 1e-4,
                  "point": "getGaussPoint2D(4, 0) = (0.861136311594053, 0.861136311594053, 0)", **d}
    i = sf.index("testShapeFunctionLagrangeQuad4")
    j = sf.index("testShapeFunctionLagrangeQuad9")
    asserts = re.findall(r"ASSERT_NEAR\((retVal4\w+),\s*([-\d.e]+)", sf[i:j])
    q4 = {}
    for name, v in asserts:
        q4[name] = float(v)
    g["quad4_at_gp4"] = {"source": "test/shape/testshapefunction.h (testShapeFunctionLagrangeQuad4)", "tol": 1e-6,
                         "values": q4}

    g["materials"] = {
        "source": "test/material/fluid/linear/elastic/testmaterialfluidideal.h:93-110; "
                  "test/material/structure/isotrop/linear/elastic/testmaterialstructureisotrop.h:91-118; "
                  "test/material/structure/isotrop/linear/viscoelastic/testmaterialstructurevisco.h:92-142",
        "fluid": {"c": 343.0, "rho": 1.21},
        "isotrop": {"E": 7e10, "nu": 0.3},
        "visco": {"E": 7e10, "eta": 0.001, "type": 1, "E_vis": [7e10, 7e7]},
    }

    h5 = os.path.join(REF, "resources/unitHDF5/material/fluid/linear/elastic/"
                           "acousticmedium_fluid8_2elem_AF_LIN_UAF_ISO_DIR.hdf5")
    txt = subprocess.run(["/opt/conda/bin/h5dump", h5], capture_output=True, text=True, check=True).stdout
    el = re.search(r'DATASET "mtxFemElemGroup1".*?DATA \{(.*?)\}', txt, re.S).group(1)
    rows = [[int(v) for v in re.findall(r"\d+", line.split(":", 1)[1])]
            for line in el.strip().splitlines() if ":" in line]
    nd = re.search(r'DATASET "mtxFemNodes".*?DATA \{(.*)\}\s*\}\s*\}\s*GROUP "Nodesets"', txt, re.S).group(1)
    recs = re.findall(r"\{\s*(\d+),\s*([-\d.e]+),\s*([-\d.e]+),\s*([-\d.e]+)\s*\}", nd)
    g["fluid8_2elem_fixture"] = {
        "source": "test/resources/unitHDF5/material/fluid/linear/elastic/"
                  "acousticmedium_fluid8_2elem_AF_LIN_UAF_ISO_DIR.hdf5 (h5dump)",
        "elements": rows,
        "nodes": [[int(r[0]), float(r[1]), float(r[2]), float(r[3])] for r in recs],
        "rho": 1.21, "c": 343.0, "fixed_pressure_node": 7,
    }
    with open(OUT, "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
