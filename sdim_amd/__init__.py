"""sdim_amd - MI355X (gfx950) engine for sdim's prime-dimension stabilizer tableau path."""
from .engine import Engine, Job, Ticket, MEASUREMENT_DTYPE, default_engine
