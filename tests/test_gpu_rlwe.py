"""The reference example's RLWE flow (example.cpp:12-58) on the device
(yell_amd.rlwe over the C ABI, samples from the device samplers) against the
oracle's flow on the same samples drawn by the host samplers with the same
seeds: the NTT-form secret, the public key, both ciphertext halves and the
decryption, bit-exact.  The encryption + decryption captured as a graph and
replayed matches the eager run on every replay."""
import numpy as np
import pytest

from _oracle import Oracle, rlwe_oracle
from gpu_util import device, say
from yell_amd import capi
from yell_amd.rlwe import Rlwe

pytestmark = pytest.mark.gpu
R = rlwe_oracle()


def host_samples(n, L, word, seeds_pk, seeds_ct, hwt):
    h = lambda kind, param, s: capi.sample_host(kind, n, L, param, s, word)
    return (h("hwt", hwt, seeds_pk[0]), h("gaussian", 3.2, seeds_pk[1]), h("uniform", None, seeds_pk[2]),
            h("zo", 0x7F, seeds_ct[0]), h("gaussian", 3.2, seeds_ct[1]), h("gaussian", 3.2, seeds_ct[2]))


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("n,L", [(8, 1), (1024, 2), (8192, 4), (16384, 8)])
def test_device_flow_matches_oracle(word, n, L):
    say("rlwe flow", word, n, L)
    device()
    hwt = min(64, n)
    f = Rlwe(n, L, word=word, hwt=hwt)
    sx, bx, ax, cb, ca, u, msg, plain = (f.buffer() for _ in range(8))
    seeds_pk, seeds_ct = (11, 12, 13), (21, 22, 23)
    msg_h = capi.sample_host("uniform", n, L, None, 99, word)
    msg.upload(msg_h)
    f.keygen(sx, bx, ax, seeds_pk)
    f.encrypt(cb, ca, u, msg, bx, ax, seeds_ct)
    f.decrypt(plain, cb, ca, sx)
    got = [b.download() for b in (sx, bx, cb, ca, plain)]
    s_sx, s_e0, s_ax, s_u, s_e1, s_e2 = host_samples(n, L, word, seeds_pk, seeds_ct, hwt)
    np.testing.assert_array_equal(ax.download(), s_ax)
    want = R.flow(Oracle(word), n, L, s_sx, s_e0, s_ax, msg_h, s_u, s_e1, s_e2)
    for name, g, w in zip(("sx", "pk_bx", "ct_bx", "ct_ax", "plain"), got, want):
        np.testing.assert_array_equal(g, w, err_msg=name)
    for b in (sx, bx, ax, cb, ca, u, msg, plain):
        b.close()


@pytest.mark.parametrize("fused", [True, False])
def test_many_messages_match_single_flow(fused):
    """encrypt_many / decrypt_many of 6 messages = encrypt / decrypt of each
    with seeds + k, bit-exact, with the steps after each sampler fused and as
    separate launches."""
    device()
    n, L, count = 8192, 4, 6
    f = Rlwe(n, L)
    sx, bx, ax = f.buffer(), f.buffer(), f.buffer()
    f.keygen(sx, bx, ax, (1, 2, 3))
    msgs = capi.DeviceBuffer(n * L * count)
    capi.call("ya_sample_uniform_device_many", msgs.ptr, n, L, count, 500, None)
    cb, ca, u, pl = (f.buffer(count) for _ in range(4))
    say("rlwe many", count, "fused", fused)
    f.encrypt_many(cb, ca, u, msgs, bx, ax, count, (10, 20, 30), fused=fused)
    f.decrypt_many(pl, cb, ca, sx, count, fused=fused)
    got = [x.download().reshape(count, n * L) for x in (cb, ca, pl)]
    w = n * L
    m_all = msgs.download()
    for k in range(count):
        m1, c1, a1, u1, p1 = (f.buffer() for _ in range(5))
        m1.upload(m_all[k * w:(k + 1) * w])
        f.encrypt(c1, a1, u1, m1, bx, ax, (10 + k, 20 + k, 30 + k))
        f.decrypt(p1, c1, a1, sx)
        for name, g, one in zip(("ct_bx", "ct_ax", "plain"), got, (c1, a1, p1)):
            np.testing.assert_array_equal(g[k], one.download(), err_msg="%s of message %d" % (name, k))
        for b in (m1, c1, a1, u1, p1):
            b.close()
    for b in (sx, bx, ax, msgs, cb, ca, u, pl):
        b.close()


def test_graph_replay_matches_eager():
    """Capture encrypt + decrypt of 4 messages on a created stream, replay it
    three times with eager work in between: each replay equals the eager run."""
    lib = device()
    n, L, B = 8192, 4, 4
    stream = capi.ctypes.c_void_p()
    capi.call("ya_stream_create", capi.ctypes.byref(stream))
    st = stream.value
    f = Rlwe(n, L, stream=st)
    sx, bx, ax = f.buffer(), f.buffer(), f.buffer()
    f.keygen(sx, bx, ax, (1, 2, 3))
    msgs = [capi.DeviceBuffer.from_host(capi.sample_host("uniform", n, L, None, 50 + k, 64)) for k in range(B)]
    work = [(f.buffer(), f.buffer(), f.buffer(), f.buffer()) for _ in range(B)]
    capi.call("ya_stream_synchronize", None)  # the uploads ran on the null stream

    def step():
        for k, (cb, ca, u, pl) in enumerate(work):
            f.encrypt(cb, ca, u, msgs[k], bx, ax, (100 + 3 * k, 101 + 3 * k, 102 + 3 * k))
            f.decrypt(pl, cb, ca, sx)

    step()  # eager (also builds the Gaussian table before capture)
    capi.call("ya_stream_synchronize", st)
    eager = [[b.download() for b in w] for w in work]
    say("rlwe graph capture")
    capi.call("ya_graph_begin", st)
    step()
    g = capi.ctypes.c_void_p()
    capi.call("ya_graph_end", st, capi.ctypes.byref(g))
    for rep in range(3):
        for w in work:  # eager work between replays: clobber the outputs
            for b in w:
                b.upload(np.zeros(n * L, dtype=np.uint64))
        capi.call("ya_stream_synchronize", None)
        capi.call("ya_graph_launch", g.value, st)
        capi.call("ya_stream_synchronize", st)
        for k, w in enumerate(work):
            for e, b in zip(eager[k], w):
                np.testing.assert_array_equal(b.download(), e, err_msg="replay %d msg %d" % (rep, k))
    capi.call("ya_graph_destroy", g.value)
    capi.call("ya_stream_destroy", st)
