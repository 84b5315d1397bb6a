from .rules import Rules
from .parser import Parser
