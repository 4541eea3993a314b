"""The launch-shape registry behind FockPlan.autotune (host logic, no device):
record / lookup / count / clear, the text file round trip, architecture
keys, rejected entries and files, and VIBEQC_TUNING_FILE persistence across
processes."""

from __future__ import annotations

import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def tuning():
    from vibeqc_amd import tuning

    tuning.clear()
    yield tuning
    tuning.clear()


def test_record_lookup_and_clear(tuning):
    assert tuning.count() == 0
    assert tuning.lookup("gfx950", "bf16x3", 1536, 320, 7424) is None
    tuning.record("gfx950", "bf16x3", 1536, 320, 7424, 5, 24, 57.5)
    tuning.record("gfx950", "fp64", 1536, 320, 7424, 0, 12, 320.0)
    assert tuning.lookup("gfx950", "bf16x3", 1536, 320, 7424) == (5, 24)
    assert tuning.lookup("gfx950", "fp64", 1536, 320, 7424) == (0, 12)
    # Every key field matters: architecture, precision, nbf, occupied, ldq.
    assert tuning.lookup("gfx942", "bf16x3", 1536, 320, 7424) is None
    assert tuning.lookup("gfx950", "fp32", 1536, 320, 7424) is None
    assert tuning.lookup("gfx950", "bf16x3", 1535, 320, 7424) is None
    assert tuning.lookup("gfx950", "bf16x3", 1536, 319, 7424) is None
    assert tuning.lookup("gfx950", "bf16x3", 1536, 320, 7552) is None
    # Recording a key again replaces its shape.
    tuning.record("gfx950", "bf16x3", 1536, 320, 7424, 10, 0, 56.0)
    assert tuning.lookup("gfx950", "bf16x3", 1536, 320, 7424) == (10, 0)
    assert tuning.count() == 2
    tuning.clear()
    assert tuning.count() == 0


@pytest.mark.parametrize(
    "entry",
    [
        ("gfx950", "bf16x3", 1536, 320, 7424, 11, 0),  # tile beyond 10 column blocks
        ("gfx950", "bf16x3", 1536, 320, 7424, -1, 0),
        ("gfx950", "bf16x3", 1536, 320, 7424, 1, 129),  # more slices than the partial buffer allows
        ("gfx950", "bf16x3", 1536, 320, 7400, 1, 1),  # ldq not a multiple of 128
        ("gfx950", "bf16x3", 320, 1536, 7424, 1, 1),  # more occupied columns than AOs
        ("gfx 950", "bf16x3", 1536, 320, 7424, 1, 1),  # blank in the architecture name
        ("", "bf16x3", 1536, 320, 7424, 1, 1),
    ],
)
def test_invalid_entries_are_refused(tuning, entry):
    from vibeqc_amd._native import NativeError

    with pytest.raises(NativeError) as err:
        tuning.record(*entry)
    assert err.value.status == 1
    assert tuning.count() == 0


def test_file_round_trip_and_merge(tuning, tmp_path):
    path = tmp_path / "jk.tuning"
    assert tuning.load(path) is False  # missing file: nothing loaded, no error
    tuning.record("gfx950", "bf16x3", 490, 81, 2432, 3, 48, 0.97)
    tuning.record("gfx942", "fp32", 384, 80, 1920, 2, 0, 1.5)
    tuning.save(path)
    lines = path.read_text().splitlines()
    assert lines[0] == "vibeqc-jk-tuning 1" and len(lines) == 3
    tuning.clear()
    tuning.record("gfx950", "fp64", 24, 5, 128, 1, 2, 0.1)
    assert tuning.load(path) is True
    assert tuning.count() == 3  # merged with what was there
    assert tuning.lookup("gfx950", "bf16x3", 490, 81, 2432) == (3, 48)
    assert tuning.lookup("gfx942", "fp32", 384, 80, 1920) == (2, 0)
    assert tuning.lookup("gfx950", "fp64", 24, 5, 128) == (1, 2)


@pytest.mark.parametrize(
    "text",
    [
        "gfx950 2 490 81 2432 3 48 0.97\n",  # no header
        "vibeqc-jk-tuning 2\ngfx950 2 490 81 2432 3 48 0.97\n",  # other version
        "vibeqc-jk-tuning 1\ngfx950 2 490 81 2432 3 48\n",  # field missing
        "vibeqc-jk-tuning 1\ngfx950 2 490 81 2432 3 48 0.97 extra\n",
        "vibeqc-jk-tuning 1\ngfx950 7 490 81 2432 3 48 0.97\n",  # unknown precision
        "vibeqc-jk-tuning 1\ngfx950 2 490 81 2400 3 48 0.97\n",  # ldq not a multiple of 128
    ],
)
def test_malformed_files_are_refused_whole(tuning, tmp_path, text):
    from vibeqc_amd._native import NativeError

    path = tmp_path / "bad.tuning"
    path.write_text(text)
    with pytest.raises(NativeError) as err:
        tuning.load(path)
    assert err.value.status == 1
    assert tuning.count() == 0  # nothing from a malformed file is kept


def test_environment_file_persists_across_processes(tmp_path):
    path = tmp_path / "env.tuning"
    env = dict(os.environ, VIBEQC_TUNING_FILE=str(path), PYTHONPATH=ROOT)
    writer = textwrap.dedent(
        """
        from vibeqc_amd import tuning
        tuning.record("gfx950", "bf16x3", 1536, 320, 7424, 5, 24, 57.5)
        """
    )
    reader = textwrap.dedent(
        """
        from vibeqc_amd import tuning
        print(tuning.count(), tuning.lookup("gfx950", "bf16x3", 1536, 320, 7424))
        """
    )
    subprocess.run([sys.executable, "-c", writer], env=env, check=True, cwd=ROOT)
    assert path.read_text().splitlines()[1].split()[:7] == ["gfx950", "2", "1536", "320", "7424", "5", "24"]
    out = subprocess.run([sys.executable, "-c", reader], env=env, check=True, cwd=ROOT, capture_output=True, text=True)
    assert out.stdout.split("\n")[0] == "1 (5, 24)"
