"""The batch configurations shipped for MI355X: keys in the tuner's format
for the three benchmark workloads, values among the compiled tile shapes and
the tuner's candidate chunks.  No GPU needed."""
import json
import re

from mun_amd import autotune as A


def test_shipped_configurations_are_well_formed():
    table = A._shipped()
    assert table, "autotune_mi355x.json missing or unreadable"
    pat = re.compile(r"^cu\d+_mem\d+G\|batch\|(\d+)x(\d+)\|fp(32|64)\|(reference|exact)\|b(\d+)\|nt(\d+)\|p(\d+)$")
    meshes = set()
    for key, rec in table.items():
        m = pat.match(key)
        assert m, key
        meshes.add(int(m.group(1)))
        prec = "fp" + m.group(3)
        assert rec["tile2"] in A.TILE2_SHAPES[prec]
        assert rec["chunk"] in (16, 32, 64)
        assert rec["group"] >= 1 and rec["lanes"] in (1, 2, 3)
        assert (rec["lanes"] - 1) * rec["group"] < int(m.group(5))  # every lane gets a group
        assert rec.get("evidence")
    assert meshes == {40, 120, 200}


def test_shipped_table_is_plain_json():
    with open(A.SHIPPED) as f:
        json.load(f)
