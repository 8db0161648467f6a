"""Worker process of the CPU oracle (test infrastructure only): reads pickled
symbfact argument tuples on stdin and answers each with a pickled
(ok, result-or-message) pair on stdout.  Started by oracle.py when
ORACLE_PROCESS=1; it never loads the GPU library."""
import pickle
import sys

from oracle import oracle


def main():
    inp, out = sys.stdin.buffer, sys.stdout.buffer
    while True:
        try:
            args = pickle.load(inp)
        except EOFError:
            return
        try:
            res = (True, oracle._symbfact(*args))
        except ValueError as e:
            res = (False, str(e))
        pickle.dump(res, out, protocol=pickle.HIGHEST_PROTOCOL)
        out.flush()


if __name__ == "__main__":
    main()
