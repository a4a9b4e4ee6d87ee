"""Configuration files through the C ABI reader (psoci_configs_read, context
free, no GPU) and the Python reader, against the reference's two readers:
readConfigsToGA keeps file order (PsociConfigs.C:336-495), readConfigs
inserts into a std::set<pair<int, string>>, so it returns the pairs ordered
by symmetry and then by occupation string with duplicates gone
(PsociConfigs.C:255-270, 283-296).  Each malformed line is reported with the
reference's return value numcfgs + 2, the 1-based line number."""
import numpy as np
import pytest

from psoci_amd import api, inputs


def write(tmp_path, text, name="configs"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def both(path, sorted=False):
    return api.read_configs_file(path, sorted=sorted), inputs.read_configs(path, sorted=sorted)


def test_file_order_round_trip(tmp_path):
    cfg = inputs.mrcisd_configs(4, 6, 4)
    cfg.title = "mrci 4x6 test list"
    p = str(tmp_path / "mrci")
    inputs.write_configs(p, cfg)
    for got in both(p):
        np.testing.assert_array_equal(got.occ, cfg.occ)
        np.testing.assert_array_equal(got.jsym, cfg.jsym)
        assert (got.nbf, got.nelec, got.ksym, got.ncfg) == (cfg.nbf, cfg.nelec, cfg.ksym, cfg.ncfg)
        assert got.title == "mrci 4x6 test list"
        assert got.nopen_max == cfg.nopen_max


def test_sorted_reader_orders_by_symmetry_then_string_and_drops_duplicates(tmp_path):
    rng = np.random.default_rng(3)
    cfg = inputs.fci_configs(6, 6)
    rows = ["".join(map(str, r)) for r in cfg.occ]
    syms = rng.integers(1, 4, len(rows)).tolist()
    pairs = list(zip(syms, rows))
    dup = [pairs[i] for i in rng.choice(len(pairs), 25, replace=False)]
    lines = pairs + dup
    order = rng.permutation(len(lines))
    text = "shuffled\n6 6 1 %d 6\n" % len(lines) + "".join("%d %s\n" % lines[i] for i in order)
    p = write(tmp_path, text)
    # the set's order, from Python's own tuple/str comparison
    expect = sorted(set(pairs))
    for got in both(p, sorted=True):
        assert got.ncfg == len(expect)
        assert [(int(j), "".join(map(str, r))) for j, r in zip(got.jsym, got.occ)] == expect
    # file order keeps everything, duplicates included
    for got in both(p):
        assert got.ncfg == len(lines)
        assert [(int(j), "".join(map(str, r))) for j, r in zip(got.jsym, got.occ)] == [lines[i] for i in order]


def test_open_shell_maximum_adjusted_down_and_crlf(tmp_path):
    text = "t\r\n4 4 1 3 4\r\n1 2200\r\n1 2110\r\n1 1210\r\n"
    for got in both(write(tmp_path, text)):
        assert got.nopen_max == 2 and got.ncfg == 3 and got.title == "t"
        np.testing.assert_array_equal(got.occ, [[2, 2, 0, 0], [2, 1, 1, 0], [1, 2, 1, 0]])


GOOD = "title\n4 4 1 3 2\n1 2200\n1 2110\n1 2020\n"

BAD = [
    ("", 1),                                            # no title
    ("title only\n", 2),                                # no header
    ("t\n4 4 1 3\n1 2200\n", 2),                        # four integers
    ("t\n4 4 x 3 2\n1 2200\n", 2),                      # not an integer
    ("t\n4 4 1 3 2\n1 2200\n\n1 2020\n", 4),            # blank line
    ("t\n4 4 1 3 2\n1 2200\nA 2110\n1 2020\n", 4),      # bad symmetry
    ("t\n4 4 1 3 2\n1 2200\n1\n1 2020\n", 4),           # no occupation
    ("t\n4 4 1 3 2\n1 2200\n1 21100\n1 2020\n", 4),     # length
    ("t\n4 4 1 3 2\n1 2200\n1 2130\n1 2020\n", 4),      # character
    ("t\n4 4 1 3 1\n1 2200\n1 2110\n1 2020\n", 4),      # open shells
    ("t\n4 4 1 3 2\n1 2200\n1 2210\n1 2020\n", 4),      # electrons
    ("t\n4 4 1 4 2\n1 2200\n1 2110\n1 2020\n", 5),      # count (last line)
    ("t\n4 4 1 2 2\n1 2200\n1 2110\n1 2020\n", 5),      # count
]


@pytest.mark.parametrize("text,line", BAD)
def test_errors_name_the_line(tmp_path, text, line):
    p = write(tmp_path, text)
    for sorted_ in (False, True):
        with pytest.raises(inputs.ConfigFileError) as e:
            api.read_configs_file(p, sorted=sorted_)
        assert e.value.line == line
        with pytest.raises(inputs.ConfigFileError) as e:
            inputs.read_configs(p, sorted=sorted_)
        assert e.value.line == line


def test_sorted_count_rule(tmp_path):
    # a duplicate dropped: the count cannot be predicted, so it is accepted
    p = write(tmp_path, "t\n4 4 1 3 2\n1 2200\n1 2200\n1 2020\n")
    for got in both(p, sorted=True):
        assert got.ncfg == 2
    for got in both(p):
        assert got.ncfg == 3
    # nothing dropped and the count differs: an error in both readers
    p = write(tmp_path, "t\n4 4 1 5 2\n1 2200\n1 2110\n1 2020\n", "c2")
    for s in (False, True):
        with pytest.raises(inputs.ConfigFileError):
            api.read_configs_file(p, sorted=s)


def test_good_file_and_missing_file(tmp_path):
    for got in both(write(tmp_path, GOOD)):
        assert got.ncfg == 3 and got.nopen_max == 2
    with pytest.raises(api.PsociError) as e:
        api.read_configs_file(str(tmp_path / "absent"))
    assert e.value.code == api.ERR_ARG and "cannot open" in str(e.value)
