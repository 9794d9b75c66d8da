"""MI355X-native counting path of MERCURY (triangle counting and the P1-P8
pattern counts) behind the C ABI of ``include/mgpm.h``.

The native library is loaded on first use of an exported name, so that
``python -m mercury_amd.build`` works before anything is built.
"""
__all__ = ["Graph", "DeviceGraph", "Batch", "Server", "Tune", "Stats", "MgpmError",
           "make_tune",
           "device_count", "PATTERN_NAMES", "TC", "P1", "P2", "P3", "P4",
           "P5", "P6", "P7", "P8"]


def __getattr__(name):
    if name in __all__ or name in ("native", "lib"):
        import importlib
        native = importlib.import_module(__name__ + ".native")
        if name == "native":
            return native
        return getattr(native, name)
    raise AttributeError(name)
