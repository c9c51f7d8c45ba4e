"""Call counter decorator (dedalus/utils/function_count.py:2-27)."""

from functools import wraps


def counts(func):
    @wraps(func)
    def wrapper(*args, **kwargs):
        wrapper.count += 1
        return func(*args, **kwargs)
    wrapper.count = 0
    return wrapper
