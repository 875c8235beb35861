"""Child process that runs calls into oracle/_ref/libfeecref.so.

TEST INFRASTRUCTURE ONLY.  Started by oracle/ref.py as
``python _refworker.py <threads>``; reads pickled (kind, kwargs) requests on
stdin and writes pickled replies on the original stdout.  File descriptor 1
is pointed at stderr first, so the reference's own printing (start-up banner,
solver log lines) cannot corrupt the replies.
"""
import os
import pickle
import struct
import sys


def main():
    threads = int(sys.argv[1])
    os.environ["OMP_NUM_THREADS"] = str(threads)
    reply = os.fdopen(os.dup(1), "wb")
    os.dup2(2, 1)
    request = sys.stdin.buffer
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from oracle import ref as refmod

    refmod._init_library()
    while True:
        head = request.read(8)
        if len(head) < 8:
            return
        (size,) = struct.unpack("<Q", head)
        kind, kw = pickle.loads(request.read(size))
        try:
            out = (True, refmod._task(kind, kw))
        except Exception as exc:  # reported to the parent
            out = (False, repr(exc))
        data = pickle.dumps(out, protocol=4)
        reply.write(struct.pack("<Q", len(data)))
        reply.write(data)
        reply.flush()


if __name__ == "__main__":
    main()
