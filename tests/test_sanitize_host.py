"""The library's host code under AddressSanitizer + UndefinedBehaviorSanitizer (leak detection on,
every finding fatal): tests/sanitize/driver.cpp runs the C++ input-file readers on the reference's
data files and on malformed files, and the FP64 table builder with PA_BATCH_HOST_ONLY (no device)
on valid and invalid batches. Built by tests/sanitize/Makefile with hipcc, sanitizers on the host
compilation only."""
from __future__ import annotations

import os
import subprocess

from conftest import GOLDEN, ROOT

HERE = os.path.join(ROOT, "tests", "sanitize")

MALFORMED = {
    "truncated.cif": "data_x\n_cell_length_a 10\n",
    "empty.cif": "",
    "no_loop.cif": "data_x\n_symmetry_space_group_name_H-M 'P 1'\n_cell_length_a 10\n_cell_length_b 10\n"
                   "_cell_length_c 10\n_cell_angle_alpha 90\n_cell_angle_beta 90\n_cell_angle_gamma 90\n",
    "ragged.cif": "loop_\n_atom_site_label\n_atom_site_type_symbol\n_atom_site_fract_x\nC1 C 0.1\nC2 C\n",
    "bad_numbers.cif": "data_x\n_symmetry_space_group_name_H-M 'P 1'\n_cell_length_a ten\n",
    "unclosed.yaml": "parameters:\n  C_: [3.4\ncutoff: 12\n",
    "no_cutoff.yaml": "parameters:\n  C_: [3.4, 0.01]\n",
    "nested.yaml": "parameters:\n  C_:\n    - [[[[[[3.4]]]]]]\ncutoff: 12\n",
    "binary.bin": "\x00\x01\x02\xff" * 64,
    "long_line.txt": "x" * 200_000 + "\n",
}


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", HERE], check=True, capture_output=True)
    files = []
    for name, text in MALFORMED.items():
        p = tmp_path / name
        p.write_text(text, encoding="latin-1")
        files.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(HERE, "_build", "driver"), GOLDEN, *files,
                        os.path.join(GOLDEN, "co2.yaml"), os.path.join(GOLDEN, "RUBTAK.cif")],
                       capture_output=True, text=True, env=env, cwd=str(tmp_path), timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "SANITIZE_DRIVER_OK" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "reader calls refused" in r.stdout
