"""The .abl compiler with the -b hip backend (openabl_amd/bin/openabl-hip), on
the CPU: front-end diagnostics against the reference's lint corpus, the
command-line contract, overload naming, constant folding and the shape of
the generated programs.

The lint corpus (test/*.abl with the expected stderr in test/*.exp) and the
example models are read from the reference checkout when it is present; the
tests that need them are skipped elsewhere.  Expected texts are compared as
the reference's test.sh compares them (diff -b: runs of blanks equal).
"""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ABLC = os.path.join(REPO, "openabl_amd", "bin", "openabl-hip")
REF = "/root/reference"
REF_TESTS = os.path.join(REF, "test")
REF_EXAMPLES = os.path.join(REF, "examples")
OWN_EXAMPLES = os.path.join(REPO, "examples")

pytestmark = pytest.mark.skipif(not os.path.exists(ABLC), reason="openabl-hip not built")


def ablc(*args, cwd=None):
    return subprocess.run([ABLC, *args], capture_output=True, text=True, cwd=cwd)


def squash(text):
    return [re.sub(r"\s+", " ", line).strip() for line in text.strip().splitlines()]


def lint_cases():
    if not os.path.isdir(REF_TESTS):
        return []
    return sorted(f[:-4] for f in os.listdir(REF_TESTS) if f.endswith(".abl"))


@pytest.mark.skipif(not os.path.isdir(REF_TESTS), reason="reference lint corpus not present")
@pytest.mark.parametrize("case", lint_cases())
def test_lint_corpus(case):
    r = ablc("--lint-only", "-i", os.path.join(REF_TESTS, case + ".abl"))
    with open(os.path.join(REF_TESTS, case + ".exp")) as f:
        want = f.read()
    assert squash(r.stderr) == squash(want)
    assert r.returncode == (1 if want.strip() else 0)


def example_models():
    out = []
    for d in (REF_EXAMPLES, OWN_EXAMPLES):
        if os.path.isdir(d):
            out += [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith(".abl")]
    return out


@pytest.mark.parametrize("path", example_models(), ids=os.path.basename)
def test_examples_lint_clean(path):
    r = ablc("--lint-only", "-i", path)
    assert r.returncode == 0 and r.stderr == ""


def model(tmp_path, text, name="m"):
    p = tmp_path / f"{name}.abl"
    p.write_text(text)
    return str(p)


CIRCLE_MIN = """
agent Point { position float2 pos; }
param int num_agents = 1000;
param float rho = 0.05;
param float r = 5;
float W = sqrt(num_agents / rho);
environment { max: float2(W) }
step move(Point in -> out) {
  float2 p = in.pos;
  for (Point nx : near(in, 2*r)) { p += (nx.pos - in.pos) * 0.001; }
  out.pos = clamp(p, float2(W));
}
void main() {
  for (int i : 0..num_agents) { add(Point { pos: random(float2(W)) }); }
  simulate(10) { move }
  save("out.json");
}
"""


def test_cli_contract(tmp_path):
    # option errors print the help and exit 1
    r = ablc()
    assert r.returncode == 1 and "Missing input file (-i or --input)" in r.stderr
    r = ablc("-i", "x.abl", "-Q")
    assert r.returncode == 1 and "Missing argument for option \"-Q\"" in r.stderr
    r = ablc("-i", "x.abl")
    assert r.returncode == 1 and "Missing backend (-b or --backend)" in r.stderr
    r = ablc("-i", "x.abl", "-P", "nonsense", "-b", "hip")
    assert r.returncode == 1 and 'Malformed parameter: Missing "="' in r.stderr
    assert ablc("-h").returncode == 0
    m = model(tmp_path, CIRCLE_MIN)
    r = ablc("-i", str(tmp_path / "missing.abl"), "-b", "hip", "-o", str(tmp_path))
    assert r.returncode == 1 and "could not be opened" in r.stderr
    # only the hip backend is built here
    r = ablc("-i", m, "-b", "c", "-o", str(tmp_path / "c"))
    assert r.returncode == 1 and 'Unknown backend "c"' in r.stderr
    r = ablc("-i", m, "-b", "hip")
    assert r.returncode == 1 and "Missing output directory (-o or --output-dir)" in r.stderr
    # -P of an unknown parameter, and of a constant that is not a param
    r = ablc("-i", m, "-b", "hip", "-o", str(tmp_path / "o"), "-P", "nope=3")
    assert r.returncode == 1 and 'Unknown parameter "nope" specified through -P on line 1' in r.stderr
    r = ablc("-i", m, "-b", "hip", "-o", str(tmp_path / "o"), "-P", "W=3")
    assert r.returncode == 1 and 'Only constants marked with "param"' in r.stderr
    # a feature the backend does not support: exit 2
    unsupported = CIRCLE_MIN.replace("float2 p = in.pos;",
                                     "float2 p = in.pos; for (int x : new int[3]) { p.x += 0.0; }")
    r = ablc("-i", model(tmp_path, unsupported, "u"), "-b", "hip", "-o", str(tmp_path / "u"))
    assert r.returncode == 2, r.stderr
    # a good model: the generated directory
    out = tmp_path / "gen"
    r = ablc("-i", m, "-b", "hip", "-o", str(out))
    assert r.returncode == 0, r.stderr
    for f in ("main.cpp", "steps.hip", "abl_gen.h", "build.sh", "run.sh"):
        assert (out / f).exists(), f
    assert os.access(out / "build.sh", os.X_OK) and os.access(out / "run.sh", os.X_OK)


def test_overload_names_and_folding(tmp_path):
    """Overloads are printed name_k, k the number of earlier signatures of
    the name (the standard library's random(float2, float2) is random_1
    after the builtin random(float, float), random(float2) is random_4,
    clamp(float2, float2) is clamp_4); folded constants print with 6
    significant digits and '.0' when they show no dot; -P substitutes."""
    out = tmp_path / "g"
    r = ablc("-i", model(tmp_path, CIRCLE_MIN), "-b", "hip", "-o", str(out))
    assert r.returncode == 0, r.stderr
    host = (out / "main.cpp").read_text()
    assert "float2 random_1(float2 min, float2 max)" in host
    assert "float2 random_4(float2 max)" in host
    assert "float2 clamp_4(float2 pos, float2 max)" in host
    assert "random_4(float2_fill((abl_float) (W)))" in host
    assert "abl_float PI = 3.14159;" in host
    assert "abl_float W = 141.421;" in host  # sqrt(1000 / 0.05) = 141.42135...
    dev = (out / "steps.hip").read_text()
    assert "static constexpr F W = (F) 141.421;" in dev
    assert "clamp_4(p, v2_fill<F>((F) (W)))" in dev
    r = ablc("-i", model(tmp_path, CIRCLE_MIN), "-b", "hip", "-o", str(out),
             "-P", "num_agents=1024", "-P", "rho=0.1024")
    assert r.returncode == 0
    host = (out / "main.cpp").read_text()
    assert "int num_agents = 1024;" in host and "abl_float W = 100.0;" in host
    # an int -P value for a float param is substituted as an int literal
    r = ablc("-i", model(tmp_path, CIRCLE_MIN), "-b", "hip", "-o", str(out), "-P", "r=3")
    assert "abl_float r = 3;" in (out / "main.cpp").read_text()
    r = ablc("-i", model(tmp_path, CIRCLE_MIN), "-b", "hip", "-o", str(out), "-P", "r=x")
    assert r.returncode == 1 and 'Value "x" provided for parameter "r" could not parsed' in r.stderr


def test_float_literal_printing(tmp_path):
    src = CIRCLE_MIN.replace("float W = sqrt(num_agents / rho);",
                             "float W = sqrt(num_agents / rho);\nfloat tiny = 0.00001;\n"
                             "float big = 1e3 * 1000.0;\nint folded = 7 / 2 + 1;")
    # 1e3 lexes as the integer 1 followed by the identifier e3: not a float
    r = ablc("--lint-only", "-i", model(tmp_path, src))
    assert r.returncode == 1 and "Parse error: syntax error, unexpected IDENTIFIER" in r.stderr
    src = src.replace("1e3 * 1000.0", "1.e3 * 1000.0")
    out = tmp_path / "f"
    r = ablc("-i", model(tmp_path, src), "-b", "hip", "-o", str(out))
    assert r.returncode == 0, r.stderr
    host = (out / "main.cpp").read_text()
    assert "abl_float tiny = 1.0e-05;" in host   # 1e-05, kept a C float literal
    assert "abl_float big = 1.0e+06;" in host
    assert "int folded = 4;" in host              # C integer division


def test_step_program_shape(tmp_path):
    """The device program of a step: the near loop over the grid with the C
    near test, the written columns, the near radius as the program computes
    it, registration under -C hip.program."""
    out = tmp_path / "s"
    r = ablc("-i", model(tmp_path, CIRCLE_MIN), "-b", "hip", "-o", str(out),
             "-C", "hip.program=circle_min", "-C", "hip.order=index")
    assert r.returncode == 0, r.stderr
    dev = (out / "steps.hip").read_text()
    assert 'registers program "circle_min"' in dev and '"circle_min",' in dev
    # the near test, on the square sum of nx.pos - centre
    m = re.search(r"const SqLen<F> (\w+) = sq_dist2<F>\(nx\.pos, (\w+)\);", dev)
    assert "c_.near(" in dev and m and f"c_.outside_sq({m.group(1)})" in dev
    assert "static __host__ __device__ double radius_move() { return (double) ((2.0 * r)); }" in dev
    # the near loop walks the stepping agent's own type around it and reads
    # float members only: the step runs on the runtime's LDS tiles
    assert '{"move", 0, 0, nullptr, 0.0, 3u, launch_move, false, 0, false, radius_move}' in dev
    assert "[&](const auto &nb_) -> bool {" in dev and "struct Step_move {" in dev
    assert "launch_gen_tiles<Step_move<F>>(L);" in dev and "launch_gen_lists" not in dev
    host = (out / "main.cpp").read_text()
    assert 'abl_gen_config(sim_, "hip.order", "index");' in host
    assert 'abl_save_json(&agents, agents_info, "out.json", (int) sizeof(abl_float));' in host
    # with -C hip.lists=1 the near loop (over the stepping agent's own type,
    # around it, float members only) runs on the runtime's tiled lists; the
    # key is the generator's, not passed on to the runtime
    out2 = tmp_path / "s2"
    r = ablc("-i", model(tmp_path, CIRCLE_MIN), "-b", "hip", "-o", str(out2),
             "-C", "hip.program=circle_min", "-C", "hip.lists=1")
    assert r.returncode == 0, r.stderr
    dev = (out2 / "steps.hip").read_text()
    assert '{"move", 0, 0, nullptr, 0.0, 3u, launch_move, false, 0, true, radius_move}' in dev
    assert "struct Step_move {" in dev and "launch_gen_lists<Step_move<F>>(L);" in dev
    assert "hip.lists" not in (out2 / "main.cpp").read_text()


def test_copy_of_input_member_is_not_written(tmp_path):
    """out.m = in.m alone leaves m's column as it is (copy-through), so a
    step that only copies the position does not move the grid."""
    src = """
agent Cell { position float2 pos; bool alive; }
environment { max: float2(10) }
step s(Cell in -> out) {
  int n = 0;
  for (Cell c : near(in, 1.5)) { if (c.alive) n += 1; }
  out.pos = in.pos;
  out.alive = n == 3;
}
void main() {
  for (int i : 0..10) { add(Cell { pos: float2(i, 0), alive: true }); }
  simulate(3) { s }
}
"""
    out = tmp_path / "c"
    assert ablc("-i", model(tmp_path, src), "-b", "hip", "-o", str(out)).returncode == 0
    dev = (out / "steps.hip").read_text()
    assert '{"s", 0, 0, nullptr, 0.0, 4u, launch_s,' in dev  # only the alive column
    # the near loop reads a bool member: even with -C hip.lists=1 the step
    # walks the grid, neither tiles nor lists
    out2 = tmp_path / "c2"
    assert ablc("-i", model(tmp_path, src), "-b", "hip", "-o", str(out2),
                "-C", "hip.lists=1").returncode == 0
    dev = (out2 / "steps.hip").read_text()
    assert "launch_s, false, 0, false, radius_s}" in dev and "struct Step_s" not in dev


def test_distance_compared_with_constant_takes_no_root(tmp_path):
    """A float local holding dist()/length() that is only compared with
    constants is kept as the float square sum and compared against exact
    bounds; one used in arithmetic, or compared with a local, keeps the
    root."""
    src = """
agent P { position float2 pos; float2 v; }
environment { max: float2(10) }
param float R = 2;
step s(P in -> out) {
  int n = 0;
  float2 acc = float2(0);
  for (P q : near(in, 3)) {
    float d = dist(in.pos, q.pos);
    if (d < R && 1 <= d) n += 1;
    float e = dist(in.pos, q.pos);
    if (e > 0.5) acc += (q.pos - in.pos) / e;
  }
  float sp = length(in.v);
  float lim = 1.5;
  if (sp > lim) out.v = in.v / sp;
  if (length(in.v) >= 2 * R) out.v = float2(0);
  out.pos = in.pos + acc * n;
}
void main() {
  for (int i : 0..10) { add(P { pos: float2(i, 0), v: float2(1, 0) }); }
  simulate(3) { s }
}
"""
    out = tmp_path / "d"
    r = ablc("-i", model(tmp_path, src), "-b", "hip", "-o", str(out))
    assert r.returncode == 0, r.stderr
    dev = (out / "steps.hip").read_text()
    # dist(in.pos, q.pos) in the near loop is the near test's square sum
    # (same squares, same order), reused rather than summed again
    m = re.search(r"const SqLen<F> (\w+) = sq_dist2<F>\(q\.pos, (\w+)\);", dev)
    assert m and f"c_.outside_sq({m.group(1)})" in dev
    assert f"const SqLen<F> d = {m.group(1)};" in dev
    assert "len_lt<F>(d, (double) (R))" in dev and "len_ge<F>(d, (double) (F) (1))" in dev
    assert f"F e = (F) sqrtf_rn({m.group(1)}.s);" in dev
    assert "F sp = length2<F>(in.v);" in dev
    assert "len_ge<F>(sq_length2<F>(in.v), (double) ((2.0 * R)))" in dev


def test_generated_main_links_step_library():
    """The generated host program names no symbol of libsteps.so (the step
    library registers its program with the runtime when it is loaded), so
    the link keeps it against the linker's --as-needed default."""
    gen = os.path.join(REPO, "openabl_amd", "gen", "golden_circle_f64")
    if not os.path.exists(os.path.join(gen, "main")):
        pytest.skip("generated programs not built")
    assert "-Wl,--no-as-needed -lsteps" in open(os.path.join(gen, "build.sh")).read()
    r = subprocess.run(["readelf", "-d", os.path.join(gen, "main")], capture_output=True, text=True)
    assert "[libsteps.so]" in r.stdout


def test_near_sum_reused_only_for_loop_and_centre_positions(tmp_path):
    """dist() of the loop agent's and the centre's positions, in either
    order, is the near test's square sum; dist() of any other pair (another
    member, a shifted position) is summed as written."""
    src = """
agent P { position float2 pos; float2 v; }
environment { max: float2(10) }
step s(P in -> out) {
  float a = 0;
  for (P q : near(in, 3)) {
    a += dist(q.pos, in.pos);
    a += dist(in.v, q.pos);
    a += dist(in.pos + float2(1, 0), q.pos);
  }
  out.v = float2(a, 0);
}
void main() {
  for (int i : 0..10) { add(P { pos: float2(i, 0), v: float2(1, 0) }); }
  simulate(3) { s }
}
"""
    out = tmp_path / "n"
    r = ablc("-i", model(tmp_path, src), "-b", "hip", "-o", str(out))
    assert r.returncode == 0, r.stderr
    dev = (out / "steps.hip").read_text()
    m = re.search(r"const SqLen<F> (\w+) = sq_dist2<F>\(q\.pos, (\w+)\);", dev)
    assert m and f"c_.outside_sq({m.group(1)})" in dev
    assert f"(F) sqrtf_rn({m.group(1)}.s)" in dev
    assert "dist2<F>(in.v, q.pos)" in dev
    assert "dist2<F>(v2_add<F>(in.pos, " in dev
