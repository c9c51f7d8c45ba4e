"""Process helpers (dedalus/funcs.py:40-90)."""

import traceback


def insert_ipython(*args, **kwargs):
    """Interactive shell hook: unavailable without IPython in this environment."""
    try:
        from IPython import embed
    except ImportError:
        return
    embed()


def signal_print_traceback(signo, frame):
    print(traceback.print_stack(frame))


def get_mercurial_changeset_id(target_dir=None):
    return "mi355x"
