"""Rule record (reference parameters/rules.py:1-6)."""


class Rules(object):
    def __init__(self, ncells: int, activation_interval: range, distance: int, periodic: bool) -> None:
        """Cell count, activation interval, neighbourhood distance and the
        periodic flag (reference parameters/rules.py:2-6)."""
        self.ncells = ncells
        self.activation_interval = activation_interval
        self.distance = distance
        self.periodic = periodic
