"""MI355X line-strength engine: host-side Python front end (ctypes over the
C-ABI library libgain_engine.so)."""
