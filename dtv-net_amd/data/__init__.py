# Callers import this package's modules and, from the same top-level name,
# packages this directory does not define (model.PDHG, model.FISTA, ...).
# Extending __path__ with every directory of this name on sys.path lets those
# resolve from a checkout placed after this one, while the modules here win.
from pkgutil import extend_path

__path__ = extend_path(__path__, __name__)
