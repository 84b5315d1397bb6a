"""Command-line entry point (reference main.py:5-9), GPU time evolution."""
from parameters import Parser
from quantum_game import QuantumGame


def main() -> None:
    """Parse the command line and run (reference main.py:5-9)."""
    QuantumGame(Parser.instance()).start()


if __name__ == "__main__":
    main()
