"""The pfcsieve command line: argument grammar and errors (CPU), and on the GPU
the factors.txt file, the final checksum line, and checkpoint/resume."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pfcsieve_amd", "pfcsieve")


def run(args, cwd, env=None, timeout=300):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([CLI, *args], cwd=cwd, env=e, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=timeout)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")


@pytest.mark.parametrize("args,msg", [
    (["-p", "5", "-P", "1e6", "-n", "101", "-N", "200"], "-! or -# or -c argument is required"),
    (["-!", "-p", "2", "-P", "1e6", "-n", "101", "-N", "200"], "out of range argument -p 2"),
    (["-!", "-p", "5", "-P", "1e6x", "-n", "101", "-N", "200"], "invalid argument -P 1e6x"),
    (["-!", "-p", "5", "-P", "1e6", "-n", "100", "-N", "200"], "out of range argument -n 100"),
    (["-!", "-p", "5", "-P", "1e6", "-n", "101", "-N", "2147483648"], "out of range argument -N"),
    (["-!", "-v", "129", "-p", "5", "-P", "1e6", "-n", "101", "-N", "200"], "out of range argument -v 129"),
])
def test_argument_errors(tmp_path, args, msg):
    r = run(args, tmp_path)
    assert r.returncode != 0
    assert msg in r.stdout


@pytest.mark.parametrize("xml,msg", [
    (None, "init_data.xml not found, using device 0."),
    ("<app_init_data>\n<gpu_device_num>3</gpu_device_num>\n</app_init_data>\n",
     "init_data.xml: using device 3."),
    ("<app_init_data><gpu_opencl_dev_index>2</gpu_opencl_dev_index></app_init_data>",
     "init_data.xml: using device 2."),
    ("<app_init_data></app_init_data>", "init_data.xml has no GPU device number, using device 0."),
])
def test_device_from_init_data(tmp_path, xml, msg):
    """The BOINC slot's init_data.xml names the device (main.cpp:252-274); the
    message is printed before the device is opened, so this runs without a GPU
    (the run itself may then fail to open the device)."""
    if xml is not None:
        (tmp_path / "init_data.xml").write_text(xml)
    r = run(["-!", "-p", "1e9", "-P", "1000001000", "-n", "101", "-N", "200"], tmp_path, timeout=120)
    assert msg in r.stdout


def test_help(tmp_path):
    r = run(["-h"], tmp_path)
    assert r.returncode != 0 and "Program usage" in r.stdout


def _oracle_file(orc, modes, pmin, pmax, nmin, nmax):
    r = orc.run(modes, pmin, pmax, nmin, nmax)
    lines = []
    for f in r.factors:
        p, nc, t = int(f["p"]), int(f["nc"]), int(f["type"])
        form = {0: "!", 1: "#", 2: "!/#"}[t]
        lines.append(f"{p} | {abs(nc)}{form}{'+1' if nc > 0 else '-1'}\n")
    tail = f"{r.checksum:016X}\n" if r.factorcount else f"no factors\n{r.checksum:016X}\n"
    return "".join(lines) + tail, r


@pytest.mark.gpu
@pytest.mark.parametrize("flags,job", [(["-!", "-c"], (5, 101, 20000, 101, 3000)),
                                       (["-#"], (2, 101, 20000, 101, 3000)),
                                       (["-!"], (1, 10**9, 10**9 + 10**5, 101, 3000))])
def test_factors_file_matches_oracle(tmp_path, orc, flags, job):
    modes, pmin, pmax, nmin, nmax = job
    r = run([*flags, "-p", str(pmin), "-P", str(pmax), "-n", str(nmin), "-N", str(nmax), "-v", "8"], tmp_path)
    assert r.returncode == 0, r.stdout[-2000:]
    want, ref = _oracle_file(orc, *job)
    assert (tmp_path / "factors.txt").read_text() == want
    assert f"factors {ref.factorcount}, prime count {ref.primecount}, checksum {ref.checksum:016X}" in r.stdout


@pytest.mark.gpu
def test_checkpoint_resume_same_output(tmp_path, orc):
    """Stop after 3 segments (a checkpoint is written), resume, finish: the
    factors.txt file equals an uninterrupted run (cl_sieve.cpp:1406-1445)."""
    job = (5, 101, 60000, 101, 2500)
    args = ["-!", "-c", "-p", "101", "-P", "60000", "-n", "101", "-N", "2500", "-v", "8"]
    env = {"PFC_SEG_CANDIDATES": "500"}
    a = run(args, tmp_path, {**env, "PFC_STOP_AFTER_SEGMENTS": "3"})
    assert a.returncode == 0 and "Stopping after 3 segments" in a.stdout, a.stdout[-2000:]
    assert (tmp_path / "stateA.ckp").exists()
    b = run(args, tmp_path, env)
    assert b.returncode == 0 and "Resuming from checkpoint" in b.stdout, b.stdout[-2000:]
    want, _ = _oracle_file(orc, *job)
    assert (tmp_path / "factors.txt").read_text() == want
    c = run(args, tmp_path, env)          # finished work unit
    assert "Workunit complete." in c.stdout


@pytest.mark.gpu
def test_corrupt_checkpoint_ignored(tmp_path, orc):
    job = (1, 10**6, 10**6 + 20000, 101, 1500)
    args = ["-!", "-p", str(10**6), "-P", str(10**6 + 20000), "-n", "101", "-N", "1500"]
    (tmp_path / "stateA.ckp").write_bytes(b"\x01" * 80)
    (tmp_path / "stateB.ckp").write_bytes(b"\x02" * 13)
    r = run(args, tmp_path)
    assert r.returncode == 0
    assert "Invalid checkpoint file stateA.ckp" in r.stdout and "Cannot parse stateB.ckp" in r.stdout
    want, _ = _oracle_file(orc, *job)
    assert (tmp_path / "factors.txt").read_text() == want


@pytest.mark.gpu
def test_cli_self_test(tmp_path):
    """`pfcsieve -s`: the reference's 16-range self-test (cl_sieve.cpp:1777-2095)
    run by the CLI on the GPU; every case must match its golden triple."""
    r = run(["-s", "-v", "8"], tmp_path, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    assert r.stdout.count("passed.") == 16
    assert "All test cases completed successfully!" in r.stdout
