"""A line that starts exactly on the right clip abscissa and ends to its right
is kept whole by the reference's signed_difference_line (src/rasterize.rs:
373-384), and its deposits past a row's end land in the next row. The GPU
flatten replaces it by the vertical lines rzg::wrap_lines makes; this CPU test
(tests/cpu_wrap.cpp) puts those lines through the oracle's deposit routine and
requires the quirk line's own deposits on every visible pixel, for dense masks,
masks into a wider image and mask_iter's (W + 1)-wide buffer."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_wrap_lines_deposit_like_the_kept_line(tmp_path):
    exe = str(tmp_path / "cpu_wrap")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-D_GNU_SOURCE", "-o", exe,
                    os.path.join(HERE, "cpu_wrap.cpp")], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert int(r.stdout.split()[2]) > 10_000  # wrapped deposits were exercised
