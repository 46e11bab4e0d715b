"""Restart rotation of a run directory (the reference's restart.py:9-55):
backups with the next numeric suffix, fin_config.xyz -> init_config.xyz, the
string-valued maximum of the suffixes, and the error exits."""
import os

import pytest

from implicit_dpd_amd import restart


def _write(d, name, text):
    with open(os.path.join(d, name), "w") as f:
        f.write(text)


def _read(d, name):
    with open(os.path.join(d, name)) as f:
        return f.read()


def test_rotation_first_and_second(tmp_path, capsys):
    d = str(tmp_path)
    _write(d, "init_config.xyz", "old init\n")
    _write(d, "fin_config.xyz", "final state\n")
    _write(d, "run_param.txt", "params\n")
    restart.main(["restart.py", d])
    names = set(os.listdir(d))
    assert {"init_config.xyz.0", "fin_config.xyz.0", "run_param.txt.0"} <= names
    assert _read(d, "init_config.xyz.0") == "old init\n"
    assert _read(d, "init_config.xyz") == "final state\n"
    assert "fin_config.xyz -> " in capsys.readouterr().out
    # the second rotation takes suffix 1 and leaves the backups alone
    _write(d, "fin_config.xyz", "later state\n")
    restart.main(["restart.py", d])
    names = set(os.listdir(d))
    assert "init_config.xyz.1" in names and "init_config.xyz.0.1" not in names
    assert _read(d, "init_config.xyz.1") == "final state\n"
    assert _read(d, "init_config.xyz") == "later state\n"


def test_suffix_maximum_is_a_string_maximum(tmp_path):
    d = str(tmp_path)
    _write(d, "fin_config.xyz", "x\n")
    _write(d, "traject.xyz.9", "nine\n")
    _write(d, "traject.xyz.10", "ten\n")
    assert restart.backup_number(d) == 10   # max('9', '10') == '9'


def test_empty_final_config_and_usage(tmp_path):
    d = str(tmp_path)
    _write(d, "fin_config.xyz", "")
    with pytest.raises(SystemExit) as e:
        restart.main(["restart.py", d])
    assert e.value.code == 1
    with pytest.raises(SystemExit) as e:
        restart.main(["restart.py"])
    assert e.value.code == 1
