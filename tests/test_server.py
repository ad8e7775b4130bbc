"""Tests of the continuous-batching job server (sun_server_*).

* the arena that carves job memory out of the reserved budget (host check, CPU);
* every job's per-iteration observables are bit-identical to the same job run
  alone through sun_generate, whatever else is resident -- jobs of different N,
  storage, precision, volume, beta and n_or, admitted over time as memory frees up;
* admission respects the budget (jobs queue, nothing over budget is resident) and
  rejects jobs that can never fit, with SUN_ERR_BUDGET;
* cancel of pending and running jobs, unsupported and invalid requests, the
  background scheduling thread with wait();
* suspend/resume: a job taken off the device into a file (running: at its next
  iteration boundary; pending: at once) and resumed on another server finishes with
  observables bit-identical to the job never suspended.
"""
import os
import subprocess

import numpy as np
import pytest

import sun_amd

HERE = os.path.dirname(os.path.abspath(__file__))


def test_arena_randomised(tmp_path):
    exe = str(tmp_path / "arena_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "host", "arena_check.cpp")],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    assert out.strip() == "OK"


# (nc, dims, beta, storage, precision, hot, n_iter, n_or, reunit_every, seed)
JOBS = [
    (3, (8, 8, 8, 8), 6.0, 12, 4, True, 5, 4, 1, 11),
    (2, (6, 6, 6, 6), 2.3, 18, 4, False, 7, 4, 1, 12),
    (4, (4, 4, 4, 8), 11.0, 18, 4, True, 4, 4, 0, 13),
    (3, (6, 8, 8, 10), 5.8, 8, 4, True, 6, 4, 1, 14),
    (3, (8, 8, 8, 8), 6.2, 12, 4, False, 3, 4, 1, 15),
    (3, (4, 6, 6, 8), 6.0, 18, 8, True, 4, 2, 1, 16),
    (2, (8, 8, 8, 8), 2.5, 18, 8, True, 5, 4, 1, 17),
    (3, (8, 8, 8, 8), 5.9, 12, 4, True, 8, 4, 1, 18),
]


def alone(ctx, job):
    nc, dims, beta, storage, prec, hot, n_iter, n_or, reunit, seed = job
    with sun_amd.Lattice(ctx, nc, dims, beta, storage, prec) as lat:
        if hot:
            lat.hot_start(seed)
        else:
            lat.cold_start()
        return lat.generate(seed=seed, n_iter=n_iter, n_or=n_or, reunit_every=reunit, first_sweep=0)


def submit(srv, job):
    nc, dims, beta, storage, prec, hot, n_iter, n_or, reunit, seed = job
    return srv.submit(nc, dims, beta, seed, n_iter, n_or=n_or, reunit_every=reunit, storage=storage,
                      precision=prec, hot_start=hot)


def job_bytes(job):
    nc, dims, beta, storage, prec = job[:5]
    return sun_amd.lattice_bytes(nc, dims, storage, prec, beta)


@pytest.mark.gpu
def test_mix_equals_alone_with_queueing(ctx):
    """A budget of about 2.5 of the largest jobs (half the total) forces queueing
    and backfill; each job's history equals its alone run bit for bit."""
    budget = 2 * max(job_bytes(j) for j in JOBS) + (512 << 10)
    assert budget < sum(job_bytes(j) for j in JOBS)
    with sun_amd.Server(ctx, budget, max_jobs=16) as srv:
        ids = [submit(srv, j) for j in JOBS]
        peak_running, steps = 0, 0
        while True:
            st = srv.stats()
            if st.jobs_done == len(JOBS):
                break
            srv.step()
            st = srv.stats()
            peak_running = max(peak_running, st.running)
            assert st.bytes_in_use <= st.budget
            steps += 1
            assert steps < 200
        assert 1 < peak_running < len(JOBS)
        assert srv.stats().bytes_in_use == 0
        for job, jid in zip(JOBS, ids):
            st = srv.status(jid)
            assert st.state == sun_amd.JOB_DONE and st.iters_done == job[6]
            assert st.t_submit <= st.t_start <= st.t_end
            got = srv.observables(jid, job[6])
            ref = alone(ctx, job)
            assert np.array_equal(got, ref), (job, np.abs(got - ref).max())
            assert got[-1, 0] == st.plaq[0] and got[-1, 3] == st.poly[1]
            srv.release(jid)
        # link updates: (1 + n_or) sweeps of 4V links per iteration
        expect = sum(j[6] * (1 + j[7]) * 4 * int(np.prod(j[1])) for j in JOBS)
        assert srv.stats().link_updates == expect


LARGE_JOBS = [  # volume 2^17: groups of large volume (single-lattice update kernels)
    (3, (16, 16, 16, 32), 6.0, 12, 4, True, 3, 4, 1, 31),
    (3, (16, 16, 16, 32), 5.8, 12, 4, False, 2, 4, 1, 32),
    (2, (16, 16, 16, 32), 2.4, 18, 4, True, 2, 2, 1, 33),
    (3, (16, 16, 32, 16), 6.1, 18, 8, True, 2, 1, 1, 34),
]


@pytest.mark.gpu
def test_large_volume_groups_equal_alone(ctx):
    """Jobs of at least 2^17 sites share a group per volume whose updates run through
    the single-lattice kernels (two SU(3) FP32-12 jobs in one group, SU(2) and FP64
    groups beside it); each history equals its alone run bit for bit."""
    budget = sum(job_bytes(j) for j in LARGE_JOBS) + (4 << 20)
    with sun_amd.Server(ctx, budget, max_jobs=8) as srv:
        ids = [submit(srv, j) for j in LARGE_JOBS]
        steps = 0
        while srv.stats().jobs_done < len(LARGE_JOBS):
            srv.step()
            steps += 1
            assert steps < 50
        for job, jid in zip(LARGE_JOBS, ids):
            got = srv.observables(jid, job[6])
            ref = alone(ctx, job)
            assert np.array_equal(got, ref), (job, np.abs(got - ref).max())
            srv.release(jid)


@pytest.mark.gpu
def test_large_volume_jobs_run_the_context_shapes(ctx):
    """A context's launch shapes for a lattice description (sun_ctx_set_launch) are
    what its large-volume server jobs of that description run; lattices created on it
    start with them; the history equals the job's alone run on a context without."""
    job = LARGE_JOBS[0]
    nc, dims, storage, prec = job[0], job[1], job[3], job[4]
    with sun_amd.Context(0, 0, 0) as c2:
        c2.set_launch(nc, dims, 6.0, storage, prec, 1, 256, (2, 2, 1), 16384, 1)
        c2.set_launch(nc, dims, 6.0, storage, prec, 0, 128, (-1, 0, 0), 0, 0)
        with sun_amd.Lattice(c2, nc, dims, 6.0, storage, prec) as lat:
            o, h = lat.get_launch(1), lat.get_launch(0)
            assert list(o.brick) == [2, 2, 1] and o.lds == 16384 and o.piped == 1 and o.block == 256
            assert h.brick[0] < 0 and h.block == 128 and h.piped == 0
        with pytest.raises(sun_amd.SunError) as e:
            c2.set_launch(nc, dims, 6.0, storage, prec, 1, 256, (5, 0, 0), 0, 0)
        assert e.value.code == sun_amd.ERR_INVALID
        with pytest.raises(sun_amd.SunError):
            c2.set_launch(2, dims, 2.4, 18, 8, 1, 256, (-1, 0, 0), 0, 2)  # no pipelined SU(2) kernel
        with sun_amd.Server(c2, job_bytes(job) + (4 << 20), max_jobs=2) as srv:
            jid = submit(srv, job)
            steps = 0
            while srv.stats().jobs_done < 1:
                srv.step()
                steps += 1
                assert steps < 20
            got = srv.observables(jid, job[6])
    assert np.array_equal(got, alone(ctx, job))


@pytest.mark.gpu
def test_arrivals_while_running(ctx):
    """Jobs submitted while others are mid-run join at the next step and still equal
    their alone runs."""
    a, b, c = JOBS[0], JOBS[1], JOBS[2]
    with sun_amd.Server(ctx, 1 << 30) as srv:
        ia = submit(srv, a)
        srv.step()
        srv.step()
        ib = submit(srv, b)
        srv.step()
        ic = submit(srv, c)
        while srv.stats().jobs_done < 3:
            srv.step()
        for job, jid in ((a, ia), (b, ib), (c, ic)):
            assert np.array_equal(srv.observables(jid, job[6]), alone(ctx, job))


@pytest.mark.gpu
def test_background_thread_and_wait(ctx):
    with sun_amd.Server(ctx, 1 << 30) as srv:
        srv.start()
        ids = [submit(srv, j) for j in JOBS[:4]]
        for jid in ids:
            srv.wait(jid, 60.0)
        srv.stop()
        for job, jid in zip(JOBS[:4], ids):
            assert srv.status(jid).state == sun_amd.JOB_DONE
            assert np.array_equal(srv.observables(jid, job[6]), alone(ctx, job))
        # step() is refused while the thread runs, allowed after stop()
        srv.start()
        with pytest.raises(sun_amd.SunError) as e:
            srv.step()
        assert e.value.code == sun_amd.ERR_STATE
        srv.stop()
        srv.step()


@pytest.mark.gpu
def test_budget_rejection_and_bad_requests(ctx):
    big = (3, (16, 16, 16, 16), 6.0, 18, 8, True, 2, 4, 1, 1)
    with sun_amd.Server(ctx, job_bytes(big) // 2) as srv:
        with pytest.raises(sun_amd.SunError) as e:
            submit(srv, big)
        assert e.value.code == sun_amd.ERR_BUDGET
        with pytest.raises(sun_amd.SunError) as e:  # cold start in 8-real storage
            srv.submit(3, (4, 4, 4, 4), 6.0, 1, 2, storage=8, precision=4, hot_start=False)
        assert e.value.code == sun_amd.ERR_UNSUPPORTED
        with pytest.raises(sun_amd.SunError) as e:
            srv.submit(3, (4, 4, 4, 4), 6.0, 1, 2, reunit_every=2)
        assert e.value.code == sun_amd.ERR_UNSUPPORTED
        with pytest.raises(sun_amd.SunError) as e:  # 12-real links need reunitarisation
            srv.submit(3, (4, 4, 4, 4), 6.0, 1, 20, reunit_every=0, storage=12, precision=4)
        assert e.value.code == sun_amd.ERR_UNSUPPORTED
        with pytest.raises(sun_amd.SunError) as e:  # odd extent
            srv.submit(3, (4, 4, 4, 5), 6.0, 1, 2)
        assert e.value.code == sun_amd.ERR_INVALID
        with pytest.raises(sun_amd.SunError) as e:
            srv.submit(3, (4, 4, 4, 4), 6.0, 1, 0)
        assert e.value.code == sun_amd.ERR_INVALID
        with pytest.raises(sun_amd.SunError) as e:
            srv.status(999)
        assert e.value.code == sun_amd.ERR_STATE
        # the server is still usable
        jid = srv.submit(3, (4, 4, 4, 4), 6.0, 5, 2)
        while srv.status(jid).state != sun_amd.JOB_DONE:
            srv.step()


@pytest.mark.gpu
def test_cancel_pending_and_running(ctx):
    j = JOBS[0]
    with sun_amd.Server(ctx, job_bytes(j) * 3 // 2 + (64 << 10)) as srv:
        first = submit(srv, j)
        second = submit(srv, j)  # does not fit beside the first: stays pending
        srv.step()
        assert srv.status(first).state == sun_amd.JOB_RUNNING
        assert srv.status(second).state == sun_amd.JOB_PENDING
        srv.cancel(second)
        assert srv.status(second).state == sun_amd.JOB_CANCELLED
        srv.cancel(first)
        srv.step()
        st = srv.status(first)
        assert st.state == sun_amd.JOB_CANCELLED and st.iters_done < j[6]
        assert srv.stats().bytes_in_use == 0 and srv.stats().running == 0
        with pytest.raises(sun_amd.SunError):
            srv.observables(first, 1)
        srv.release(first)
        srv.release(second)
        third = submit(srv, j)
        while srv.status(third).state != sun_amd.JOB_DONE:
            srv.step()
        assert np.array_equal(srv.observables(third, j[6]), alone(ctx, j))


def run_until(srv, jid, state, max_steps=200):
    for _ in range(max_steps):
        if srv.status(jid).state == state:
            return
        srv.step()
    raise AssertionError(f"job {jid} did not reach state {state}")


@pytest.mark.gpu
def test_suspend_running_and_resume_elsewhere(ctx, tmp_path):
    job = JOBS[7]  # 8 iterations
    path = str(tmp_path / "job.bin")
    with sun_amd.Server(ctx, 1 << 30) as srv:
        other = submit(srv, JOBS[1])
        jid = submit(srv, job)
        srv.step()
        srv.step()
        srv.suspend(jid, path)
        srv.step()  # honoured at this iteration boundary
        st = srv.status(jid)
        assert st.state == sun_amd.JOB_SUSPENDED and st.iters_done == 3
        with pytest.raises(sun_amd.SunError):
            srv.observables(jid, 1)
        run_until(srv, other, sun_amd.JOB_DONE)
        assert np.array_equal(srv.observables(other, JOBS[1][6]), alone(ctx, JOBS[1]))
        assert srv.stats().bytes_in_use == 0
    with sun_amd.Server(ctx, 1 << 30) as srv2:
        j2 = srv2.resume(path)
        assert srv2.status(j2).iters_done == 3
        run_until(srv2, j2, sun_amd.JOB_DONE)
        assert np.array_equal(srv2.observables(j2, job[6]), alone(ctx, job))


@pytest.mark.gpu
def test_suspend_pending_and_resumed_pending(ctx, tmp_path):
    j = JOBS[0]
    p1, p2, p3 = (str(tmp_path / f"j{i}.bin") for i in range(3))
    with sun_amd.Server(ctx, job_bytes(j) * 3 // 2 + (64 << 10)) as srv:
        first = submit(srv, j)
        second = submit(srv, j)  # no room beside the first: pending
        srv.step()
        srv.suspend(second, p1)  # pending: written at once, no lattice
        assert srv.status(second).state == sun_amd.JOB_SUSPENDED
        srv.step()
        srv.suspend(first, p2)
        srv.step()
        assert srv.status(first).state == sun_amd.JOB_SUSPENDED
        # both resumed; the second one waits for memory, and is suspended again while
        # pending (its lattice travels from the first file to the third)
        a = srv.resume(p2)
        b = srv.resume(p1)
        srv.step()
        c = srv.resume(p2)
        assert srv.status(c).state == sun_amd.JOB_PENDING
        srv.suspend(c, p3)
        run_until(srv, a, sun_amd.JOB_DONE)
        run_until(srv, b, sun_amd.JOB_DONE)
        ref = alone(ctx, j)
        assert np.array_equal(srv.observables(a, j[6]), ref)
        assert np.array_equal(srv.observables(b, j[6]), ref)
        d = srv.resume(p3)
        run_until(srv, d, sun_amd.JOB_DONE)
        assert np.array_equal(srv.observables(d, j[6]), ref)
        with pytest.raises(sun_amd.SunError):
            srv.resume(str(tmp_path / "missing.bin"))
        open(p1, "r+b").write(b"XXXXXXXX")
        with pytest.raises(sun_amd.SunError, match="job file"):
            srv.resume(p1)


@pytest.mark.gpu
def test_suspend_with_background_thread(ctx, tmp_path):
    # long enough that the suspend request lands before the job finishes
    job = (3, (8, 8, 8, 8), 6.0, 12, 4, True, 60, 4, 1, 19)
    path = str(tmp_path / "bg.bin")
    with sun_amd.Server(ctx, 1 << 30) as srv:
        srv.start()
        jid = submit(srv, job)
        srv.suspend(jid, path)
        srv.wait(jid, 60.0)
        assert srv.status(jid).state == sun_amd.JOB_SUSPENDED
        j2 = srv.resume(path)
        srv.wait(j2, 60.0)
        srv.stop()
        assert np.array_equal(srv.observables(j2, job[6]), alone(ctx, job))


def job_file(path, magic=b"SUNJOB01", version=1, has_lattice=0, nc=3, dims=(8, 8, 8, 8), beta=6.0, storage=12,
             prec=4, seed=1, hot=1, n_iter=10, n_or=4, reunit=1, iters_done=0, rows=0):
    """a job file in the server's layout (magic, version, has-lattice flag, the job
    description field by field, iterations done, observable rows, checkpoint)"""
    import struct
    with open(path, "wb") as f:
        f.write(struct.pack("<8sIIi4idiiQiiiii", magic, version, has_lattice, nc, *dims, beta, storage, prec, seed,
                            hot, n_iter, n_or, reunit, iters_done))
        f.write(b"\0" * (32 * rows))
    return path


@pytest.mark.gpu
def test_resume_rejects_bad_job_files(ctx, tmp_path):
    """Corrupt or inconsistent job files are refused with a status, before any
    allocation sized from the file; a well-formed pending job file is accepted."""
    p = str(tmp_path / "j.bin")
    cases = [
        (dict(magic=b"NOTAJOB!"), sun_amd.ERR_INVALID),
        (dict(version=2), sun_amd.ERR_INVALID),
        (dict(n_iter=0), sun_amd.ERR_INVALID),
        (dict(n_or=-1), sun_amd.ERR_INVALID),
        (dict(reunit=2), sun_amd.ERR_UNSUPPORTED),
        (dict(nc=5), sun_amd.ERR_UNSUPPORTED),
        (dict(dims=(8, 8, 7, 8)), sun_amd.ERR_INVALID),
        (dict(iters_done=11), sun_amd.ERR_INVALID),
        (dict(iters_done=3), sun_amd.ERR_INVALID),                        # progress without a lattice
        (dict(has_lattice=1, iters_done=3, rows=1), sun_amd.ERR_INVALID),  # truncated rows
        (dict(has_lattice=1, iters_done=3, rows=3), sun_amd.ERR_INVALID),  # no checkpoint after them
        (dict(has_lattice=1, n_iter=2 ** 31 - 1, iters_done=2 ** 30), sun_amd.ERR_BUDGET),
        (dict(storage=8, hot=0), sun_amd.ERR_UNSUPPORTED),
        (dict(storage=12, reunit=0), sun_amd.ERR_UNSUPPORTED),
        # a finished job (no iteration left): launching it would write past its history
        (dict(has_lattice=1, n_iter=3, iters_done=3, rows=3), sun_amd.ERR_INVALID),
    ]
    with sun_amd.Server(ctx, 1 << 30) as srv:
        for fields, code in cases:
            with pytest.raises(sun_amd.SunError) as e:
                srv.resume(job_file(p, **fields))
            assert e.value.code == code, (fields, e.value.code, str(e.value))
        assert srv.stats().pending == 0 and srv.stats().bytes_in_use == 0
        jid = srv.resume(job_file(p, n_iter=2, seed=9))
        run_until(srv, jid, sun_amd.JOB_DONE)
        assert srv.status(jid).iters_done == 2


@pytest.mark.gpu
def test_suspend_to_unwritable_path_keeps_jobs_running(ctx, tmp_path):
    """A suspend request whose file cannot be created is refused at the call; the
    job and its neighbours run on and finish with their alone results."""
    a, b = JOBS[7], JOBS[0]
    with sun_amd.Server(ctx, 1 << 30) as srv:
        ja, jb = submit(srv, a), submit(srv, b)
        srv.step()
        with pytest.raises(sun_amd.SunError) as e:
            srv.suspend(ja, str(tmp_path / "no" / "such" / "dir" / "job.bin"))
        assert e.value.code == sun_amd.ERR_INVALID
        run_until(srv, ja, sun_amd.JOB_DONE)
        run_until(srv, jb, sun_amd.JOB_DONE)
        assert srv.job_error(ja) == ""
        assert np.array_equal(srv.observables(ja, a[6]), alone(ctx, a))
        assert np.array_equal(srv.observables(jb, b[6]), alone(ctx, b))


def resumable_job_file(ctx, path, U, job, tmp_path):
    """a job file (no iterations done) whose lattice is the host field U"""
    nc, dims, beta, storage, prec, hot, n_iter, n_or, reunit, seed = job
    ck = str(tmp_path / "field.ckpt")
    with sun_amd.Lattice(ctx, nc, dims, beta, storage, prec) as lat:
        lat.from_host(U)
        lat.save(ck, sun_amd.FILE_NATIVE, seed=seed, next_sweep=0)
    job_file(path, has_lattice=1, nc=nc, dims=dims, beta=beta, storage=storage, prec=prec, seed=seed,
             hot=int(hot), n_iter=n_iter, n_or=n_or, reunit=reunit)
    with open(path, "ab") as f:
        f.write(open(ck, "rb").read())
    return path


@pytest.mark.gpu
def test_observable_guard_fails_only_the_bad_job(ctx, tmp_path):
    """A job whose field is not finite (set from the host, resumed from a job file)
    or far from SU(N) (every link doubled: plaquette 16) ends in JOB_FAILED with the
    iteration and values in its message, while a job beside it in the same group
    finishes with its alone result; sun_generate reports such a field as an error."""
    good = (3, (8, 8, 8, 8), 6.0, 18, 4, True, 6, 4, 0, 31)  # the same group as the bad jobs
    nc, dims = good[0], good[1]
    with sun_amd.Lattice(ctx, nc, dims, 6.0, 18, 4) as lat:
        lat.hot_start(5)
        U = lat.to_host()
    nan_field = U.copy()
    nan_field[2, 17] = np.nan
    big_field = np.zeros_like(U)  # twice the unit matrix: every plaquette is 16
    big_field[..., np.arange(nc), np.arange(nc)] = 2.0
    bad_nan = (3, dims, 6.0, 18, 4, True, 6, 4, 0, 41)
    bad_big = (3, dims, 6.0, 18, 4, True, 6, 4, 0, 42)
    p1 = resumable_job_file(ctx, str(tmp_path / "nan.bin"), nan_field, bad_nan, tmp_path)
    p2 = resumable_job_file(ctx, str(tmp_path / "big.bin"), big_field, bad_big, tmp_path)
    with sun_amd.Server(ctx, 1 << 30) as srv:
        jg = submit(srv, good)
        j1, j2 = srv.resume(p1), srv.resume(p2)
        run_until(srv, jg, sun_amd.JOB_DONE)
        for j in (j1, j2):
            run_until(srv, j, sun_amd.JOB_FAILED)
            assert srv.status(j).iters_done >= 1
            with pytest.raises(sun_amd.SunError):
                srv.observables(j, 1)
        assert "not finite or outside [-1, 1]" in srv.job_error(j1) and "nan" in srv.job_error(j1)
        assert "iteration 0" in srv.job_error(j2), srv.job_error(j2)
        assert srv.job_error(jg) == ""
        assert np.array_equal(srv.observables(jg, good[6]), alone(ctx, good))
        assert srv.stats().bytes_in_use == 0 and srv.stats().jobs_done == 1
    with sun_amd.Lattice(ctx, nc, dims, 6.0, 18, 4) as lat:
        lat.from_host(nan_field)
        with pytest.raises(sun_amd.SunError) as e:
            lat.generate(seed=1, n_iter=2, n_or=1, reunit_every=1)
        assert e.value.code == sun_amd.ERR_STATE and "iteration 0" in str(e.value)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 12, 4) as lat:
        lat.cold_start()
        with pytest.raises(sun_amd.SunError) as e:
            lat.generate(seed=1, n_iter=2, n_or=1, reunit_every=0)
        assert e.value.code == sun_amd.ERR_UNSUPPORTED
