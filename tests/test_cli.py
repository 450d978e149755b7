"""The job-file driver asend_run_job on the CPU: its reader of the
reference's YAML subset (S/model_input.rs:34-119 line rules), without device
work (--parse-only), and the committed C1 job fixture against its generator."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "asend3d_amd", "asend_run_job")
FIXTURE = os.path.join(ROOT, "tests", "golden", "c1_job")


def parse(job, cwd):
    r = subprocess.run([CLI, job, "--parse-only"], cwd=cwd, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    counts = {}
    for line in r.stdout.splitlines():
        if line.startswith("commands "):
            it = line.split()
            counts = {it[i]: int(it[i + 1]) for i in range(0, len(it), 2)}
    return counts, r.stdout


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(CLI):
        pytest.skip("asend_run_job is not built (python -m asend3d_amd.build)")
    return CLI


def test_c1_fixture_parses(built, tmp_path):
    shutil.copytree(FIXTURE, tmp_path / "job")
    counts, out = parse("c1_job.yaml", tmp_path / "job")
    # 21 x 5 x 5 nodes, 20 x 4 x 4 elements; named sets plus one per label
    # plus "all"; three clamps, one traction, nine design variables, two terms
    assert counts == {"commands": 9, "nodes": 525, "elements": 320, "node_sets": 2 + 525 + 1,
                      "element_sets": 1 + 320 + 1, "sections": 1, "materials": 1, "constraints": 3, "loads": 1,
                      "design_variables": 9, "objective_terms": 2, "temperature_constraints": 0,
                      "thermal_loads": 0}
    assert "command writeObjective results/objectiveResults.csv" in out


def test_fixture_matches_generator(tmp_path):
    # the committed fixture is what tests/golden/make_c1_job.py writes
    gen = os.path.join(ROOT, "tests", "golden", "make_c1_job.py")
    src = open(gen).read()
    src = src.replace('OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_job")',
                      f'OUT = {str(tmp_path)!r}')
    src = src.replace("ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))",
                      f"ROOT = {ROOT!r}")
    script = tmp_path / "gen.py"
    script.write_text(src)
    subprocess.run([sys.executable, str(script)], check=True, timeout=120)
    for name in os.listdir(FIXTURE):
        assert (tmp_path / name).read_text() == open(os.path.join(FIXTURE, name)).read(), name


def test_reader_rules(built, tmp_path):
    # key levels by column, '#' comments, range lists [a, b, step), a single
    # modulus standing for all three, a later constraint file replacing the
    # model's constraints and a load file without loads leaving them
    (tmp_path / "m.yaml").write_text("""\
nodes:   # coordinates
- [0, 0.0, 0.0, 0.0]
- [1, 1.0, 0.0, 0.0]
- [2, 1.0, 1.0, 0.0]
- [3, 0.0, 1.0, 0.0]
- [4, 0.0, 0.0, 1.0]
- [5, 1.0, 0.0, 1.0]
- [6, 1.0, 1.0, 1.0]
- [7, 0.0, 1.0, 1.0]
elements:
- type: brick8
  connectivity:
  - [0, 0, 1, 2, 3, 4, 5, 6, 7]
sets:
  node:
    evens:
    - [0, 8, 2]
    base:
    - 0
    - 1
    - 2   # trailing comment
    - 3
  element:
    one:
    - 0
sections:
- type: solid
  material: m1
  elementSet: one
materials:
  m1:
    density: 1.0
    elastic:
      E: 5.0
      nu: 0.25
constraints:
- type: displacement
  terms:
  - nodeSet: base
    dof: 1
    coef: 1.0
  rhs: 0.0
loads:
- type: nodalForce
  load: [1.0, 0.0, 0.0]
  nodeSet: '6'
""")
    (tmp_path / "c.yaml").write_text("""\
constraints:
- type: displacement
  terms:
  - nodeSet: base
    dof: 1
  rhs: 0.0
- type: displacement
  terms:
  - nodeSet: evens
    dof: 3
  rhs: 0.0
""")
    (tmp_path / "l.yaml").write_text("loads:\n")
    (tmp_path / "job.yaml").write_text("""\
jobCommands:
- command: readModelInput
  fileName: m.yaml
- command: readConstraints
  fileName: c.yaml
- command: readLoads
  fileName: l.yaml
""")
    counts, _ = parse("job.yaml", tmp_path)
    assert counts["nodes"] == 8 and counts["elements"] == 1
    assert counts["node_sets"] == 2 + 8 + 1 and counts["element_sets"] == 1 + 1 + 1
    assert counts["constraints"] == 2 and counts["loads"] == 1 and counts["materials"] == 1


def test_missing_job_file_fails(built, tmp_path):
    r = subprocess.run([CLI, str(tmp_path / "none.yaml"), "--parse-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "could not open job file" in r.stderr
