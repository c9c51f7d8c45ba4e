"""Startup tasks run on ``import dedalus.mods`` (dedalus/startup_tasks.py)."""
