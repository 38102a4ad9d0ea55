"""CPU: the host C ABI exports every entry point mdk_host.h declares, and the job-table
farm (mdkh_run_table) expands a table the way the reference's farm does
(src/mpidp.cpp:198-246 table reading, :796-822 substitution, :877-897 replace_pattern)."""
import ctypes as C
import os
import re

import pytest

import megadock_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_every_declared_host_symbol_is_exported():
    text = open(os.path.join(ROOT, "include", "mdk_host.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mdkh_\w+)\s*\(", text, re.M)))
    assert "mdkh_run_table" in names
    lib = C.CDLL(M.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(M.HOST_SYMBOLS) == names


def _dry(tmp_path, table_text, extra=()):
    tb = tmp_path / "jobs.table"
    tb.write_text(table_text)
    log = tmp_path / "master.log"
    M.run_table(["-tb", str(tb), "-lg", str(log), "-dry"] + list(extra))
    text = log.read_text()
    jobs = [l.split("\t", 1) for l in text.splitlines() if re.match(r"^\d{5}\t", l)]
    return text, jobs


def test_table_expansion(tmp_path):
    text, jobs = _dry(tmp_path, "TITLE=sample jobs\nPARAM=-R $1 -L $2 -o $3.out\n"
                                "a_r.pdb\ta_l.pdb\tout/a\n\nb_r.pdb\tb_l.pdb\tout/b\n",
                      ["-N", "5", "-rt", "2", "-ot", "3"])
    assert "TITLE=sample jobs" in text and "PARAM=-R $1 -L $2 -o $3.out" in text
    assert "Retry option  : -rt 2" in text and "Output option : -ot 3" in text
    assert jobs == [["00001", "megadock-hip -N 5 -R a_r.pdb -L a_l.pdb -o out/a.out"],
                    ["00002", "megadock-hip -N 5 -R b_r.pdb -L b_l.pdb -o out/b.out"]]


def test_table_keywords_and_leading_spaces(tmp_path):
    # spaces among the first 7 characters are dropped (src/mpidp.cpp:211, :251-275), so
    # " param=" is a PARAM line; every occurrence of $n is replaced; repeated spaces
    # between tokens give no empty arguments
    text, jobs = _dry(tmp_path, "  title=t\n param=-R $1  -L $1 -t 3\nx.pdb\n")
    assert "TITLE=t" in text
    assert jobs == [["00001", "megadock-hip -R x.pdb -L x.pdb -t 3"]]


def test_missing_table_is_an_error(tmp_path):
    with pytest.raises(M.MdkError, match="was not opened"):
        M.run_table(["-tb", str(tmp_path / "none.table"), "-lg", str(tmp_path / "l.log"), "-dry"])
    with pytest.raises(M.MdkError, match="no job table"):
        M.run_table(["-lg", str(tmp_path / "l.log")])


def test_specialised_size_lists_match_the_build():
    """The Python lists of specialised sizes are the ones the library is built with."""
    mk = open(os.path.join(ROOT, "megadock_amd", "csrc", "Makefile")).read().replace("\\\n", " ")
    sizes = [int(x) for x in re.search(r"^PIPE_SIZES := (.*)$", mk, re.M).group(1).split()]
    assert sizes == M.FURTHER_NF
    src = open(os.path.join(ROOT, "megadock_amd", "csrc", "mdk_pipe.hip")).read()
    listed = [int(x) for x in re.findall(r"X\((\d+)\)", src)]
    assert listed == M.FURTHER_NF
    assert all("case %d: return run_pipe<%d, 2>" % (n, n) in src for n in M.TUNED_NF)


def test_command_line_parses_afresh_after_a_trailing_flag(tmp_path):
    """A parse that ends on a flag (-O) must not leak into the next one: the second call
    reads its own arguments (here it reaches the GPU check, or fails on the missing PDB)."""
    for args in (["-R", "r.pdb", "-L", "l.pdb", "-O"], ["-R", str(tmp_path / "none_r.pdb"), "-L", "l.pdb", "-t", "2"]):
        with pytest.raises(M.MdkError) as e:
            M.run(args)
        assert "invalid option" not in str(e.value) and "Usage" not in str(e.value), str(e.value)
