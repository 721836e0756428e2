"""Host-side lifetime rules of the Python binding (no device needed): a
captured graph keeps the device buffers it writes alive."""
import gc

import riggle_amd


class FakeLib:
    """Stands in for libchuckle_hip.so: hands out fake device addresses and
    records which of them were freed."""

    def __init__(self):
        self.next = 0x1000
        self.freed = []
        self.graphs_freed = 0

    def chk_dev_alloc(self, idx, nbytes, out):
        out._obj.value = self.next
        self.next += 0x1000
        return 0

    def chk_dev_free(self, idx, ptr):
        self.freed.append(ptr.value)
        return 0

    def chk_graph_create(self, *args):
        args[-1]._obj.value = 0xBEEF
        return 0

    def chk_graph_free(self, h):
        self.graphs_freed += 1
        return 0

    def chk_last_error(self):
        return b""


class FakeIndex:
    handle = 0x10


def test_graph_keeps_captured_buffers_alive(monkeypatch):
    fake = FakeLib()
    monkeypatch.setattr(riggle_amd, "_lib", fake)
    ix = FakeIndex()
    q = FakeIndex()
    g = riggle_amd.Graph(ix, q, riggle_amd.DeviceBuffer(ix, 64), riggle_amd.DeviceBuffer(ix, 64),
                         riggle_amd.DeviceBuffer(ix, 64), riggle_amd.DeviceBuffer(ix, 64),
                         riggle_amd.DeviceBuffer(ix, 8))
    gc.collect()
    assert fake.freed == [], "temporaries passed to Graph were freed while it can still launch"
    assert all(b.ptr for b in g._bufs if b is not None)
    g.close()
    gc.collect()
    assert fake.graphs_freed == 1
    assert len(fake.freed) == 5, "buffers are released once the graph is gone"
