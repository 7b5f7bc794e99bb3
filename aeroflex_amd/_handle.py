"""Owner of one C-ABI handle (a solver, mesh, batch, mapper, lattice, dense
factorisation, beam or coupling), with explicit release.

Every handle class derives from Handle and names its destroy function.  A
handle is released by close(), by leaving a `with` block, or, for handles the
caller never closed, by close_all(), which runs from Python's atexit at the
start of interpreter shutdown.  Releasing from there, rather than from
__del__ while modules are being torn down, keeps every hipFree, stream and
event destruction ahead of the HIP runtime's own exit handlers and of any
profiler that finalises at process exit.

close_all() releases live handles newest first: a handle that refers to other
handles (a batch to its solvers, a mapper to its two solvers, a coupling to
its lattice and beam) is always created after them, so it goes first.
"""
import atexit
import itertools
import threading
import weakref

_live = {}                      # creation number -> weak reference
_counter = itertools.count()
_lock = threading.Lock()


class Handle:
    """Base of the handle classes.  Subclasses set self.h through _own() and
    implement _destroy(h)."""

    h = None

    def _own(self, h):
        self.h = h
        with _lock:
            self._seq = next(_counter)
            _live[self._seq] = weakref.ref(self)

    def _destroy(self, h):   # pragma: no cover - every subclass overrides it
        raise NotImplementedError

    def close(self):
        """Release the handle's device memory, streams and library handles
        now.  Idempotent; the object is unusable afterwards."""
        h = self.h
        if h is None:
            return
        self.h = None
        with _lock:
            _live.pop(getattr(self, "_seq", None), None)
        self._destroy(h)

    @property
    def closed(self):
        return self.h is None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def live_handles():
    """Number of handles not yet released."""
    with _lock:
        return sum(1 for r in _live.values() if r() is not None)


def close_all():
    """Release every live handle, newest first."""
    with _lock:
        refs = [r for _, r in sorted(_live.items(), reverse=True)]
    for r in refs:
        obj = r()
        if obj is not None:
            obj.close()


atexit.register(close_all)
