from .gridBoundary import GridBoundary

__all__ = ["GridBoundary"]
