"""Write integrator_coefficients.json: the literal coefficients and stage
order of the reference's 3- and 4-stage integrators, read as text from
hmclab/Samplers.py (_propagate_4_stage_simplified, lines 1586-1592, and
_propagate_3_stage_simplified, lines 1663-1669).  Nothing is imported or run
from the reference.  Usage: python tests/golden/make_integrator_coefficients.py
[path to Samplers.py]."""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/hmclab/Samplers.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "integrator_coefficients.json")


def block(text, name):
    i = text.index(f"def {name}(")
    j = text.index("\n    def ", i + 1)
    return text[i:j]


def literals(body):
    """Coefficient lines `a1 = 0.0713...` (plain decimal literals) and the
    derived ones (`a3 = 1.0 - 2.0 * a1 - 2.0 * a2`), as written."""
    lit, derived = {}, {}
    for m in re.finditer(r"^\s+([ab]\d) = ([^\n]+)$", body, flags=re.M):
        name, rhs = m.group(1), m.group(2).strip()
        if name in lit or name in derived:
            continue
        if re.fullmatch(r"[0-9.]+", rhs):
            lit[name] = rhs
        else:
            derived[name] = rhs
    schema = re.search(r"# Schema: \(([^)]*)\)", body).group(1).split(",")
    # the order the loop body applies them: "# A1", "# B1", ... comments
    order = [c.lower() for c in re.findall(r"^\s+# ([AB]\d)\s*$", body, flags=re.M)]
    return {"literal": lit, "derived": derived, "schema": schema, "loop_order": order}


text = open(SRC).read()
out = {
    "source": "hmclab/Samplers.py: _propagate_4_stage_simplified (1586-1661), "
              "_propagate_3_stage_simplified (1663-1726)",
    "4s": literals(block(text, "_propagate_4_stage_simplified")),
    "3s": literals(block(text, "_propagate_3_stage_simplified")),
}
with open(OUT, "w") as f:
    json.dump(out, f, indent=1)
print(json.dumps(out, indent=1))
