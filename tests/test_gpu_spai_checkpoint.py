"""Checkpoints of the learned SPAI model (spk_spai_model_save / _load; no
reference counterpart -- the reference's only resume is the warm-start x):
the file holds the weights, Adam's moments and step, so a loaded model
predicts the same M and continues training bit for bit like the one that
was saved; damaged or foreign files are rejected and leave the model as it
was."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 6345
SIZE = 8 + 4 + 4 + 8 + 3 * NP * 4 + 8


def _mats(spk, k=6, side=24):
    return [spk.aniso_q1_2d(side, s, 1.5) for s in range(k)]


def test_checkpoint_resumes_training_exactly(spk, tmp_path):
    mats = _mats(spk)
    a = spk.SpaiModel(3)
    a.train(mats, epochs=2, learning_rate=2e-3, shuffle_seed=5)
    path = str(tmp_path / "model.spai")
    a.save(path)
    b = spk.SpaiModel(99).load(path)
    assert np.array_equal(a.params, b.params)
    A = spk.aniso_q1_2d(40, 777, 1.5)
    ra, ca, va = a.predict(A).download()
    rb, cb, vb = b.predict(A).download()
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb) and np.array_equal(va, vb)
    # both continue with the same optimiser state
    a.train(mats, epochs=2, learning_rate=2e-3, shuffle_seed=6)
    b.train(mats, epochs=2, learning_rate=2e-3, shuffle_seed=6)
    assert np.array_equal(a.params, b.params)
    # control: the weights alone (optimiser reset) continue differently
    c = spk.SpaiModel(0)
    c.load(path)
    c.params = c.params
    c.train(mats, epochs=2, learning_rate=2e-3, shuffle_seed=6)
    assert not np.array_equal(a.params, c.params)


def test_checkpoint_layout(spk, tmp_path):
    mats = _mats(spk, k=3)
    m = spk.SpaiModel(1)
    m.train(mats, epochs=2, learning_rate=1e-3, batch_matrices=2, shuffle_seed=2)
    path = tmp_path / "m.spai"
    m.save(str(path))
    raw = path.read_bytes()
    assert len(raw) == SIZE
    assert raw[:8] == b"SPKSPAI\0"
    ver, npar, step = struct.unpack_from("<IIq", raw, 8)
    assert (ver, npar, step) == (1, NP, 2 * 2)  # 2 epochs x ceil(3 / 2) steps
    w = np.frombuffer(raw, np.float32, NP, 24)
    assert np.array_equal(w, m.params)
    h = 0xcbf29ce484222325
    for byte in raw[:-8]:
        h = ((h ^ byte) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    assert struct.unpack_from("<Q", raw, len(raw) - 8)[0] == h


@pytest.mark.parametrize("damage", ["flip", "truncate", "magic", "missing"])
def test_checkpoint_rejects_damaged_files(spk, tmp_path, damage):
    src = spk.SpaiModel(4)
    src.train(_mats(spk, k=2), epochs=1, learning_rate=1e-3)
    path = tmp_path / "m.spai"
    src.save(str(path))
    raw = bytearray(path.read_bytes())
    if damage == "flip":
        raw[1000] ^= 0x10
        path.write_bytes(bytes(raw))
    elif damage == "truncate":
        path.write_bytes(bytes(raw[:-100]))
    elif damage == "magic":
        raw[0:8] = b"NOTSPAI\0"
        path.write_bytes(bytes(raw))
    else:
        path = tmp_path / "absent.spai"
    dst = spk.SpaiModel(8)
    before = dst.params
    with pytest.raises(spk.ReadError):
        dst.load(str(path))
    assert np.array_equal(dst.params, before)
