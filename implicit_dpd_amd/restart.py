"""Restart rotation of a run directory, the reference's `restart.py`
(restart.py:9-55) in Python 3.

`python -m implicit_dpd_amd.restart ROOT_DIR`:

1. backup (restart.py:27-40): every entry of ROOT_DIR whose name does not
   contain a backup suffix `word.word.digits` is copied to `name.N`, where N
   is one more than the largest suffix among the existing backups, or 0 when
   there are none.  As in the reference, the largest suffix is the largest
   *string* (restart.py:23: `max` over the matched digit strings), so with
   backups `.9` and `.10` the next one is `.10` again.
2. init config (restart.py:42-47): `fin_config.xyz` is copied over
   `init_config.xyz`; an empty `fin_config.xyz` stops with status 1.

Each copy is logged as `src -> dst`, and a wrong argument count prints the
usage and exits with status 1 (restart.py:49-52).
"""
import os
import re
import shutil
import sys

BACKUP = r'(\w+)\.(\w+)\.(\d+)'


def copy_with_log(src, dst):
    print(f"{src} -> {dst}")
    shutil.copy(src, dst)


def backup_number(root_dir, pattern=BACKUP):
    nums = set()
    for f in os.listdir(root_dir):
        m = re.search(pattern, f)
        if m is not None:
            nums.add(m.group(3))
    return int(max(nums)) + 1 if nums else 0


def make_backup(root_dir):
    num = backup_number(root_dir)
    for f in sorted(os.listdir(root_dir)):
        if re.search(BACKUP, f) is not None:  # already a backup
            continue
        copy_with_log(os.path.join(root_dir, f), os.path.join(root_dir, f"{f}.{num}"))


def make_init_config(root_dir):
    init_config = os.path.join(root_dir, "init_config.xyz")
    fin_config = os.path.join(root_dir, "fin_config.xyz")
    if os.path.getsize(fin_config) == 0:
        print(f"WARNING! there is no trajectory data in {fin_config}")
        sys.exit(1)
    copy_with_log(fin_config, init_config)


def main(argv):
    if len(argv) != 2:
        print(f"Usage: {argv[0]} root_dir")
        sys.exit(1)
    make_backup(argv[1])
    make_init_config(argv[1])


if __name__ == "__main__":
    main(sys.argv)
