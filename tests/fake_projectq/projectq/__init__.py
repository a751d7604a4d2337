from .cengines import MainEngine  # noqa: F401
