from dataclasses import field


def param(default=None, *, unit: str = "", description: str = "", factory=None):
    """dataclass field carrying the reference's ``unit``/``description`` metadata."""
    meta = {"unit": unit, "description": description}
    if factory is not None:
        return field(default_factory=factory, metadata=meta)
    return field(default=default, metadata=meta)
