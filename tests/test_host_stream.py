"""Frame streams (nbody_data_stream / nbody_data_stream_reader) and the run
loop's dumps, on the CPU: the reference's stream test
(test/stream/test_nbody_stream.cpp:62-240) with the oracle's `simple` engine
standing for nbody_engine_simple, then its failure branches one by one."""
import os

import numpy as np

import oracle_lib
from nbody_amd import nbody_create_solver, nbody_data, nbody_data_stream, nbody_data_stream_reader


def _universe(stars=128, box=100.0):
    d = nbody_data()
    d.set_quiet(True)
    d.make_universe(stars, box, box, box)
    return d


def _count_lines(path):
    with open(path) as f:
        return len([line for line in f.read().split("\n") if line])


def test_run_writes_frames_and_reader_seeks(tmp_path):
    """test_nbody_stream::run: euler, dt 0.1, dumps every 0.1 to t = 0.31 with
    two frames per data file; the reader sees 5 frames over 3 files, times and
    steps as written, and the last frame equals the engine's final state."""
    data = _universe()
    e = oracle_lib.simple_engine(1)
    assert e.init(data)
    s = nbody_create_solver({"solver": "euler"})
    s.set_engine(e)
    s.set_time_step(0.01, 0.1)
    base = str(tmp_path / "stream-test" / "new")
    stream = nbody_data_stream()
    assert stream.open(base, 14000) == 0
    assert s.run(data, 0.31, 0.1, 0.1, stream=stream) == 0
    stream.close()
    for k in range(3):
        assert os.path.exists(f"{base}{k}.dat")
    assert not os.path.exists(f"{base}3.dat")
    assert _count_lines(base + ".idx") == 8
    with open(base + ".idx") as f:
        lines = f.read().split("\n")
    assert lines[:3] == ["#coord_size 8", f"#body_count {data.get_count()}", "#box_size 100"]
    assert lines[3].split("\t") == ["0", "0", "0", "0"]
    # a time is written as QTextStream writes a double (6 significant digits):
    # the frame at 0.1 + 0.1 + 0.1 = 0.30000000000000004 reads back as 0.3
    assert lines[6].split("\t")[:2] == ["3", "0.3"]

    reader = nbody_data_stream_reader()
    got, expected = nbody_data(), data.copy()
    got.resize(data.get_count())
    assert reader.load(base) == 0
    assert reader.get_frame_count() == 5
    assert reader.get_steps_count() == 4
    assert reader.get_max_time() == 0.4
    assert reader.get_body_count() == data.get_count()
    assert reader.get_coord_size() == 8
    assert reader.get_last_file_n() == 2
    e.get_data(expected)

    assert reader.seek(0) == 0
    assert reader.seek(77) != 0
    assert reader.seek(0) == 0
    assert reader.read(got) == 0
    assert not got.is_equal(expected, 0.0)
    assert reader.seek(reader.get_frame_count() - 1) == 0
    assert reader.read(got) == 0
    assert got.is_equal(expected, 0.0)
    # the run loop leaves the last frame in `data`
    assert data.is_equal(expected, 0.0)
    assert reader.seek(0) == 0
    for frame in range(5):
        assert reader.read(got) == 0
        assert got.get_step() == frame
        assert got.get_time() == [0.0, 0.1, 0.2, 0.3, 0.4][frame]
        assert got.is_equal(expected, 0.0) == (frame == 4)
    # masses and colours come back from the .mass and .col files
    _, _, m, _ = got.bodies()
    assert np.array_equal(m, data.bodies()[2])
    assert np.array_equal(got.colors(), data.colors())
    want = np.array([[0.0, 1.0, 0.0, 1.0], [0.4, 0.4, 0.8, 1.0], [0.8, 0.4, 0.4, 1.0]], np.float32)
    assert set(map(tuple, data.colors())) == set(map(tuple, want))   # black holes and the two galaxies
    s.destroy()
    e.destroy()


def test_resume_continues_the_stream(tmp_path):
    """Resume as the reference CLI does (test/main/main.cpp:36-55): load the
    stream, read its last frame, continue it in append mode; the frames and
    the final state equal those of one uninterrupted run."""
    full, part = str(tmp_path / "full" / "s"), str(tmp_path / "part" / "s")
    states = {}
    for name, base, t1 in (("full", full, 0.6), ("part", part, 0.3)):
        data = _universe(64)
        e = oracle_lib.simple_engine(1)
        assert e.init(data)
        s = nbody_create_solver({"solver": "rk4", "max_step": 0.05})
        s.set_engine(e)
        st = nbody_data_stream()
        assert st.open(base, 1 << 30) == 0
        assert s.run(data, t1, 0.1, 0.0, stream=st) == 0
        st.close()
        states[name] = data
        s.destroy()
        e.destroy()
    reader = nbody_data_stream_reader()
    assert reader.load(part) == 0
    assert reader.seek(reader.get_frame_count() - 1) == 0
    data = nbody_data()
    data.set_quiet(True)
    data.resize(reader.get_body_count())
    assert reader.read(data) == 0
    st = nbody_data_stream()
    assert st.open(part, 1 << 30, append_to=reader) == 0
    reader.close()
    e = oracle_lib.simple_engine(1)
    assert e.init(data)
    s = nbody_create_solver({"solver": "rk4", "max_step": 0.05})
    s.set_engine(e)
    assert s.run(data, 0.6, 0.1, 0.0, stream=st) == 0
    st.close()
    assert data.is_equal(states["full"], 0.0)
    with open(full + ".idx") as a, open(part + ".idx") as b:
        assert a.read() == b.read()
    with open(full + "0.dat", "rb") as a, open(part + "0.dat", "rb") as b:
        assert a.read() == b.read()
    s.destroy()
    e.destroy()


def test_negative_branches(tmp_path):
    """test_nbody_stream::negative_branch, in the same order."""
    small = _universe(32, 1.0)
    st = nbody_data_stream()
    assert st.write(None) != 0
    st = nbody_data_stream()
    assert st.open("/dev/null/:nameless", 1000) != 0
    assert st.write(small) != 0
    st = nbody_data_stream()
    assert st.write(small) != 0                  # never opened
    st = nbody_data_stream()
    assert st.open(str(tmp_path / "00" / "stream"), 1000) == 0
    assert st.open(str(tmp_path / "01" / "stream"), 1000) == 0
    assert st.write(small) == 0
    # a directory where the data file or the index should be
    os.makedirs(tmp_path / "0" / "stream0.dat")
    st = nbody_data_stream()
    assert st.open(str(tmp_path / "0" / "stream"), 1) != 0
    assert st.write(small) != 0
    os.makedirs(tmp_path / "1" / "stream.idx")
    st = nbody_data_stream()
    assert st.open(str(tmp_path / "1" / "stream"), 1) != 0
    assert st.write(small) != 0
    # the second data file cannot be created
    os.makedirs(tmp_path / "2" / "stream1.dat")
    st = nbody_data_stream()
    assert st.open(str(tmp_path / "2" / "stream"), 1) == 0
    assert st.write(small) == 0
    assert st.write(small) != 0

    r = nbody_data_stream_reader()
    assert r.load("/dev/null/:nameless") != 0
    for k, text in ((3, "0\t1\t2\n"), (4, "a\tb\tc\td\n")):
        os.makedirs(tmp_path / str(k))
        (tmp_path / str(k) / "stream.idx").write_text(text)
        assert nbody_data_stream_reader().load(str(tmp_path / str(k) / "stream")) != 0
    r = nbody_data_stream_reader()
    assert r.get_frame_count() == 0 and r.get_steps_count() == 0 and r.get_max_time() == 0

    data = _universe()
    # the data file of frame 0 is gone: load fails (it seeks to frame 0)
    st = nbody_data_stream()
    assert st.open(str(tmp_path / "5" / "stream"), 1) == 0
    assert st.write(data) == 0 and st.write(data) == 0
    st.close()
    os.remove(tmp_path / "5" / "stream0.dat")
    assert nbody_data_stream_reader().load(str(tmp_path / "5" / "stream")) != 0
    # the data file of frame 1 is gone: frame 0 reads, the next read fails
    st = nbody_data_stream()
    assert st.open(str(tmp_path / "6" / "stream"), 1) == 0
    assert st.write(data) == 0 and st.write(data) == 0
    st.close()
    os.remove(tmp_path / "6" / "stream1.dat")
    r = nbody_data_stream_reader()
    d = nbody_data()
    d.resize(data.get_count())
    assert r.load(str(tmp_path / "6" / "stream")) == 0
    assert r.read(d) == 0
    assert r.read(d) != 0

    # index edited: negative offsets, wrong coordinate size, zero body count
    st = nbody_data_stream()
    assert st.open(str(tmp_path / "7" / "stream"), 1) == 0
    for _ in range(3):
        assert st.write(data) == 0
    st.close()
    idx = tmp_path / "7" / "stream.idx"
    idx.write_text("#coord_size 8\n#body_count 256\n#box_size 100\n1\t1\t0\t0\n1\t1\t1\t-1\n1\t1\t2\t-1\n")
    r = nbody_data_stream_reader()
    d = nbody_data()
    d.resize(data.get_count())
    assert r.load(str(tmp_path / "7" / "stream")) == 0
    assert r.read(d) == 0
    assert r.read(d) != 0
    assert r.read(d) != 0
    idx.write_text("#coord_size 6\n#body_count 256\n1\t1\t0\t0\n")
    assert nbody_data_stream_reader().load(str(tmp_path / "7" / "stream")) != 0
    idx.write_text("#coord_size 8\n#body_count 0\n#box_size 100\n1\t1\t0\t0\n")
    assert nbody_data_stream_reader().load(str(tmp_path / "7" / "stream")) != 0
    # a reader for another body count
    idx.write_text("#coord_size 8\n#body_count 256\n#box_size 100\n1\t1\t0\t0\n")
    r = nbody_data_stream_reader()
    assert r.load(str(tmp_path / "7" / "stream")) == 0
    d = nbody_data()
    d.resize(10)
    assert r.read(d) != 0
    assert r.read(None) != 0
